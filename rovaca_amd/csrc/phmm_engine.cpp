// Host side of the MI355X PairHMM engine: device set-up, packing of regions
// into the device format, length-binned tile queues, launches and the C ABI
// of include/rovaca_phmm.h.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <emmintrin.h>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rovaca_phmm.h"
#include "phmm_adapter.h"
#include "phmm_engine.h"
#include "phmm_internal.h"
#include "phmm_tables.h"

namespace rvc {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace {
thread_local std::string t_last_error;
}

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_last_error = buf;
  return code;
}
const char* last_error() { return t_last_error.c_str(); }

#define RVC_HIP(call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return set_error(RVC_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// engine-wide state
// ---------------------------------------------------------------------------
Engine g_engine;

int engine_init(const rvc_config* cfg) {
  std::lock_guard<std::mutex> lk(g_engine.mu);
  if (g_engine.ready) return set_error(RVC_E_STATE, "rvc_phmm_init called twice");
  const int dev = cfg ? cfg->device : 0;
  int n = 0;
  RVC_HIP(hipGetDeviceCount(&n));
  if (dev < 0 || dev >= n) return set_error(RVC_E_INVALID, "device %d out of range (%d devices)", dev, n);
  RVC_HIP(hipSetDevice(dev));
  hipDeviceProp_t prop;
  RVC_HIP(hipGetDeviceProperties(&prop, dev));
  g_engine.device = dev;
  g_engine.n_cu = prop.multiProcessorCount;
  auto* ht = new (std::nothrow) HostTables;
  if (!ht) return set_error(RVC_E_INVALID, "out of host memory");
  build_tables(ht);
  std::memcpy(ht->gop_cap, pcr_gop_caps(), 3 * 21);
  g_engine.host_tables.reset(ht);
  void* p = nullptr;
  const size_t bytes = sizeof(HostTables);
  RVC_HIP(hipMalloc(&p, bytes));
  RVC_HIP(hipMemcpy(p, ht, bytes, hipMemcpyHostToDevice));
  g_engine.dev_tables_mem = p;
  auto* base = static_cast<HostTables*>(p);
  g_engine.tables.ph2pr_f = base->ph2pr_f;
  g_engine.tables.mm_f = base->mm_f;
  g_engine.tables.ph2pr_d = base->ph2pr_d;
  g_engine.tables.mm_d = base->mm_d;
  g_engine.tables.gop_cap = base->gop_cap;
  g_engine.hbm_budget = (cfg && cfg->hbm_budget_bytes) ? cfg->hbm_budget_bytes : (64ull << 30);
  g_engine.max_batch_regions = (cfg && cfg->max_batch_regions) ? cfg->max_batch_regions : 4096u;
  g_engine.host_prep = cfg ? std::min<uint32_t>(cfg->host_prep, 2) : 0;
  g_engine.sw_min_launch = cfg ? cfg->sw_min_launch_pairs : 0;
  g_engine.sw_bt_budget = (cfg && cfg->sw_bt_budget_mib) ? (uint64_t)cfg->sw_bt_budget_mib << 20 : 12ull << 30;
  g_engine.sw_wide = cfg ? std::min<uint32_t>(cfg->sw_wide_shapes, 2) : 0;
  // kernel-shape table: the one named in the config, else tuned/<device>.txt beside the library
  std::string table;
  if (cfg && cfg->tune_table) {
    table = cfg->tune_table;
  } else {
    Dl_info info;
    if (dladdr((void*)&engine_init, &info) && info.dli_fname) {
      std::string lib = info.dli_fname;
      const size_t slash = lib.rfind('/');
      table = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/tuned/mi355x.txt";
      FILE* f = fopen(table.c_str(), "r");
      if (f) fclose(f); else table.clear();
    }
  }
  default_policy() = FP32Policy();
  g_engine.tune_source.clear();
  if (!table.empty()) {
    const int rc = load_tune_table(table.c_str(), &default_policy());
    if (rc) return rc;
    g_engine.tune_source = table;
  }
  g_engine.ready = true;
  return RVC_OK;
}

int engine_shutdown() {
  std::lock_guard<std::mutex> lk(g_engine.mu);
  if (!g_engine.ready) return set_error(RVC_E_STATE, "engine not initialised");
  if (g_engine.dev_tables_mem) hipFree(g_engine.dev_tables_mem);
  g_engine.dev_tables_mem = nullptr;
  g_engine.ready = false;
  return RVC_OK;
}

// ---------------------------------------------------------------------------
// validation and packing
// ---------------------------------------------------------------------------
namespace {
// 1 for the bytes the device format holds (A C G T, and N as the wildcard).
struct AcgtnTable {
  uint8_t ok[256] = {};
  AcgtnTable() { ok['A'] = ok['C'] = ok['G'] = ok['T'] = ok['N'] = 1; }
};
const AcgtnTable kAcgtn;

// Index of the first byte of s[0..n) outside A, C, G, T, N, or -1 (16 bytes
// per step against the five letters, then byte by byte).
int64_t first_other_base(const uint8_t* s, int32_t n) {
  int32_t x = 0;
  const __m128i a = _mm_set1_epi8('A'), c = _mm_set1_epi8('C'), g = _mm_set1_epi8('G'), t = _mm_set1_epi8('T'),
                nn = _mm_set1_epi8('N');
  for (; x + 16 <= n; x += 16) {
    const __m128i v = _mm_loadu_si128(reinterpret_cast<const __m128i*>(s + x));
    const __m128i ok = _mm_or_si128(_mm_or_si128(_mm_or_si128(_mm_cmpeq_epi8(v, a), _mm_cmpeq_epi8(v, c)),
                                                 _mm_or_si128(_mm_cmpeq_epi8(v, g), _mm_cmpeq_epi8(v, t))),
                                    _mm_cmpeq_epi8(v, nn));
    const int m = _mm_movemask_epi8(ok);
    if (m != 0xFFFF) return x + __builtin_ctz(~m & 0xFFFF);
  }
  for (; x < n; ++x)
    if (!kAcgtn.ok[s[x]]) return x;
  return -1;
}
}  // namespace

int validate_region(const rvc_region* r) {
  if (!r) return set_error(RVC_E_INVALID, "null region");
  if (r->n_haps == 0 || r->n_reads == 0) return set_error(RVC_E_INVALID, "region without haplotypes or reads");
  if (!r->hap_off || !r->read_off || !r->q || !r->i || !r->d || !r->c)
    return set_error(RVC_E_INVALID, "null array in region");
  if (r->n_haps > (uint32_t)kMaxHaps)
    return set_error(RVC_E_LIMIT, "%u haplotypes exceed the limit of %d", r->n_haps, kMaxHaps);
  for (uint32_t h = 0; h < r->n_haps; ++h) {
    const int32_t l = r->hap_off[h + 1] - r->hap_off[h];
    if (l < 0) return set_error(RVC_E_INVALID, "haplotype %u has negative length", h);
    if (l > kMaxHapLen) return set_error(RVC_E_LIMIT, "haplotype %u has %d bases (limit %d)", h, l, kMaxHapLen);
    if (l > 0 && !r->hap_bases) return set_error(RVC_E_INVALID, "null hap_bases");
  }
  for (uint32_t x = 0; x < r->n_reads; ++x) {
    const int32_t l = r->read_off[x + 1] - r->read_off[x];
    if (l < 0) return set_error(RVC_E_INVALID, "read %u has negative length", x);
    if (l > kMaxReadLen) return set_error(RVC_E_LIMIT, "read %u has %d bases (limit %d)", x, l, kMaxReadLen);
    if (l > 0 && !r->read_bases) return set_error(RVC_E_INVALID, "null read_bases");
  }
  // Bytes other than A, C, G, T, N have no single meaning in the reference
  // (raw-byte compare in its inter kernel, ph/roDV_inter/pairhmm_inter.cpp:167;
  // 'A' or out of table in its intra kernel, ph/roDV_common/common.h:42-56;
  // the kernel follows its grouping): refused, the caller keeps its CPU path.
  for (uint32_t h = 0; h < r->n_haps; ++h) {
    const int64_t x = first_other_base(r->hap_bases + r->hap_off[h], r->hap_off[h + 1] - r->hap_off[h]);
    if (x >= 0)
      return set_error(RVC_E_BASES, "haplotype %u base %lld is byte %u, not A, C, G, T or N", h, (long long)x,
                       (unsigned)r->hap_bases[r->hap_off[h] + x]);
  }
  for (uint32_t y = 0; y < r->n_reads; ++y) {
    const int64_t x = first_other_base(r->read_bases + r->read_off[y], r->read_off[y + 1] - r->read_off[y]);
    if (x >= 0)
      return set_error(RVC_E_BASES, "read %u base %lld is byte %u, not A, C, G, T or N", y, (long long)x,
                       (unsigned)r->read_bases[r->read_off[y] + x]);
  }
  return RVC_OK;
}

namespace {
inline uint32_t base_code(uint8_t ch, bool* is_n) {
  switch (ch) {
    case 'A': *is_n = false; return 0;
    case 'C': *is_n = false; return 1;
    case 'G': *is_n = false; return 2;
    case 'T': *is_n = false; return 3;
    default: *is_n = true; return 0;  // 'N' (validate_region refuses every other byte)
  }
}
inline uint32_t align_up(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }
}  // namespace

FP32Policy latency_policy(const FP32Policy& base, uint64_t pairs) {
  FP32Policy p = base;
  if (p.force_w == 0) {
    if (pairs <= 512) p.force_w = 64;
    else if (pairs <= 2048) p.force_w = 32;
  }
  return p;
}

FP32Policy& default_policy() {
  static FP32Policy p;
  return p;
}

// Table format (text): lines "occ16 <17 ints>", "occpt16 <17 ints>",
// "occpt32 <17 ints>", "occpt8 <17 ints>" (waves per SIMD), "shape16 <17
// ints>" (0 or W*256+RPL, RPL 0 = fewest rows of width W),
// "steps3 0|1", "fold 0|1" (scaled forms allowed), "table 0|1", "hap_run <1..64>"; '#' starts a
// comment.
int load_tune_table(const char* path, FP32Policy* pol) {
  FILE* f = fopen(path, "r");
  if (!f) return set_error(RVC_E_INVALID, "cannot open tune table %s", path);
  FP32Policy p;
  char line[1024];
  int rc = RVC_OK;
  while (fgets(line, sizeof line, f)) {
    char key[32];
    int n = 0;
    if (line[0] == '#' || sscanf(line, "%31s%n", key, &n) != 1) continue;
    const char* rest = line + n;
    int v[17], cnt = 0, used = 0;
    while (cnt < 17 && sscanf(rest, "%d%n", &v[cnt], &used) == 1) { rest += used; ++cnt; }
    if (!strcmp(key, "occ16") && cnt == 17) {
      for (int x = 0; x < 17; ++x) p.occ16[x] = (uint8_t)std::min(4, std::max(2, v[x]));
    } else if (!strcmp(key, "occpt16") && cnt == 17) {
      for (int x = 0; x < 17; ++x) p.occpt16[x] = (uint8_t)std::min(4, std::max(2, v[x]));
    } else if (!strcmp(key, "occpt32") && cnt == 17) {
      for (int x = 0; x < 17; ++x) p.occpt32[x] = (uint8_t)std::min(4, std::max(2, v[x]));
    } else if (!strcmp(key, "table") && cnt == 1) {
      p.prior_table = v[0] != 0;
    } else if (!strcmp(key, "occpt8") && cnt == 17) {
      for (int x = 0; x < 17; ++x) p.occpt8[x] = (uint8_t)std::min(4, std::max(2, v[x]));
    } else if (!strcmp(key, "shape16") && cnt == 17) {
      for (int x = 0; x < 17; ++x) {
        const int w = v[x] >> 8, r = v[x] & 0xff;
        // RPL 0: the fewest rows of width w for each read of class x; the
        // shortest must have a kernel (longer ones without one keep the
        // 16-lane default, FP32Policy::pick)
        const int rr = r ? r : ((x - 1) * 16 + 1 + w - 1) / std::max(1, w);
        const bool ok = v[x] == 0 || ((w == 8 || w == 16 || w == 32 || w == 64) && rr >= 1 &&
                                      (r == 0 || rr * w >= x * 16) && fp32_supported(w, rr));
        if (!ok) { rc = set_error(RVC_E_INVALID, "tune table %s: bad shape16 entry %d", path, v[x]); break; }
        p.rpl16_override[x] = (uint16_t)v[x];
      }
    } else if (!strcmp(key, "steps3") && cnt == 1) {
      p.rpl_steps_of_3 = v[0] != 0;
    } else if (!strcmp(key, "fold") && cnt == 1) {
      p.fold = v[0] != 0;
    } else if (!strcmp(key, "hap_run") && cnt == 1) {
      p.hap_run = (uint32_t)std::min<int>((int)kMaxHapRun, std::max(1, v[0]));
    } else {
      rc = set_error(RVC_E_INVALID, "tune table %s: cannot read line '%s'", path, key);
    }
    if (rc) break;
  }
  fclose(f);
  if (rc == RVC_OK) *pol = p;
  return rc;
}

void FP32Policy::pick(uint32_t R, int* W, int* RPL, bool cmp) const {
  // rows for R + 1: the segment always ends in a padding row, whose I holds
  // M + I of the read's last row (the prior-table kernel sums it directly)
  const uint32_t Rp = R + 1;
  // 16-lane shapes come in steps of 3 rows per lane, so that reads of nearby
  // lengths share one launch (each launch has a tail of its own)
  static const int kRpl16[17] = {0, 1, 2, 4, 4, 7, 7, 7, 10, 10, 10, 13, 13, 13, 16, 16, 16};
  const int exact = (int)std::max<uint32_t>(1, (Rp + 15) / 16);
  if (Rp <= 256) { *W = 16; *RPL = rpl_steps_of_3 ? kRpl16[exact] : exact; }
  else if (Rp <= 512) { *W = 32; *RPL = (int)((Rp + 31) / 32); }
  else { *W = 64; *RPL = (int)((Rp + 63) / 64); }
  bool tuned = false;
  if (Rp <= 256 && force_w) {
    *W = force_w;
    *RPL = (int)std::max<uint32_t>(1, (Rp + force_w - 1) / force_w);
  } else if (Rp <= 256 && rpl16_override[exact] != 0) {
    tuned = true;
    *W = rpl16_override[exact] >> 8;
    *RPL = rpl16_override[exact] & 0xff;
    if (*RPL == 0) {  // fewest rows of width W, in steps of 3 like the 16-lane shapes
      *RPL = (int)((Rp + *W - 1) / *W);
      if (rpl_steps_of_3 && *RPL <= 16) *RPL = kRpl16[*RPL];
    }
  }
  // a shape no kernel has (8 lanes past 19 rows) or, chosen by the tuned
  // table, one past 16 rows for the compare kernel (it spills there) keeps
  // the 16-lane default
  if (Rp <= 256 && (!fp32_supported(*W, *RPL) || (tuned && cmp && *RPL > 16))) {
    *W = 16;
    *RPL = rpl_steps_of_3 ? kRpl16[exact] : exact;
  }
}

// Appends the packed bases of seq[0..n) at base offset off (a multiple of
// kSeqAlign) of the pack / N-mask word arrays.
namespace {
// 2-bit code (bits 0-1) and N flag (bit 2) of every byte, from base_code.
struct CodeTable {
  uint8_t c[256];
  CodeTable() {
    for (int x = 0; x < 256; ++x) {
      bool is_n;
      const uint32_t code = base_code((uint8_t)x, &is_n);
      c[x] = (uint8_t)(code | (is_n ? 4u : 0u));
    }
  }
};
const CodeTable kCodes;
}  // namespace

namespace {
inline uint64_t load8(const uint8_t* p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  return v;
}
// The 2-bit codes of 8 bases (one per byte, first base in the low byte)
// packed into 16 bits, base i at bits 2i: ((c >> 1) ^ (c >> 2)) & 3 is 0, 1,
// 2, 3 for 'A', 'C', 'G', 'T' (0 for 'N'), then three shift-or steps close
// the gaps between the bytes' fields.
inline uint32_t codes8(uint64_t v) {
  uint64_t x = ((v >> 1) ^ (v >> 2)) & 0x0303030303030303ull;
  x = (x | (x >> 6)) & 0x000F000F000F000Full;
  x = (x | (x >> 12)) & 0x000000FF000000FFull;
  return (uint32_t)((x | (x >> 24)) & 0xFFFFull);
}
// Bit i set where byte i is 'N': an exact zero-byte test of v ^ "NNNNNNNN"
// (high bit of each byte), gathered the same way.
inline uint32_t nbits8(uint64_t v) {
  const uint64_t t = v ^ 0x4E4E4E4E4E4E4E4Eull;
  uint64_t z = ~(((t & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | t | 0x7F7F7F7F7F7F7F7Full) >> 7;
  z = (z | (z >> 7)) & 0x0003000300030003ull;
  z = (z | (z >> 14)) & 0x0000000F0000000Full;
  return (uint32_t)((z | (z >> 28)) & 0xFFull);
}
}  // namespace

// Widens [lo, hi] to the gap-open qualities (insertion and deletion) of one
// read and lowers clo to its smallest gap-continuation quality (7-bit
// values), 16 bases per step.
static void gap_ranges(const uint8_t* gi, const uint8_t* gd, const uint8_t* gc, uint32_t n, uint8_t* lo, uint8_t* hi,
                       uint8_t* clo) {
  const __m128i m7 = _mm_set1_epi8(127);
  __m128i vlo = _mm_set1_epi8((char)*lo), vhi = _mm_set1_epi8((char)*hi), vc = _mm_set1_epi8((char)*clo);
  uint32_t y = 0;
  for (; y + 16 <= n; y += 16) {
    const __m128i a = _mm_and_si128(_mm_loadu_si128(reinterpret_cast<const __m128i*>(gi + y)), m7);
    const __m128i b = _mm_and_si128(_mm_loadu_si128(reinterpret_cast<const __m128i*>(gd + y)), m7);
    const __m128i c = _mm_and_si128(_mm_loadu_si128(reinterpret_cast<const __m128i*>(gc + y)), m7);
    vlo = _mm_min_epu8(vlo, _mm_min_epu8(a, b));
    vhi = _mm_max_epu8(vhi, _mm_max_epu8(a, b));
    vc = _mm_min_epu8(vc, c);
  }
  alignas(16) uint8_t l[16], h[16], q[16];
  _mm_store_si128(reinterpret_cast<__m128i*>(l), vlo);
  _mm_store_si128(reinterpret_cast<__m128i*>(h), vhi);
  _mm_store_si128(reinterpret_cast<__m128i*>(q), vc);
  uint8_t L = l[0], Hh = h[0], C = q[0];
  for (int k = 1; k < 16; ++k) {
    L = std::min(L, l[k]);
    Hh = std::max(Hh, h[k]);
    C = std::min(C, q[k]);
  }
  for (; y < n; ++y) {
    const uint8_t a = gi[y] & 127, b = gd[y] & 127, c = gc[y] & 127;
    L = std::min(L, std::min(a, b));
    Hh = std::max(Hh, std::max(a, b));
    C = std::min(C, c);
  }
  *lo = L;
  *hi = Hh;
  *clo = C;
}

static void pack_seq(const uint8_t* seq, uint32_t n, uint32_t off, std::vector<uint32_t>& pack,
                     std::vector<uint32_t>& nmask) {
  // off is a multiple of 32: whole 2-bit and N-mask words, assembled in registers
  uint32_t* pw = pack.data() + (off >> 4);
  uint32_t* nw = nmask.data() + (off >> 5);
  uint32_t x = 0;
  for (; x + 32 <= n; x += 32) {
    const uint64_t v0 = load8(seq + x), v1 = load8(seq + x + 8), v2 = load8(seq + x + 16), v3 = load8(seq + x + 24);
    pw[x >> 4] |= codes8(v0) | (codes8(v1) << 16);
    pw[(x >> 4) + 1] |= codes8(v2) | (codes8(v3) << 16);
    nw[x >> 5] |= nbits8(v0) | (nbits8(v1) << 8) | (nbits8(v2) << 16) | (nbits8(v3) << 24);
  }
  for (; x < n; ++x) {  // the last, partial word
    const uint32_t c = kCodes.c[seq[x]];
    pw[x >> 4] |= (c & 3u) << ((x & 15u) * 2u);
    nw[x >> 5] |= (c >> 2) << (x & 31u);
  }
}

namespace {
// Per-thread working storage of the packing steps, kept from call to call.
struct PackScratch {
  std::vector<std::vector<Tile>> bins = std::vector<std::vector<Tile>>(kFp32Bins);
  std::vector<std::vector<uint32_t>> bin_cost = std::vector<std::vector<uint32_t>>(kFp32Bins);
  std::vector<uint8_t> hap_has_n;
  std::vector<uint32_t> order, idx;
  std::vector<Tile> out, bin;
  std::vector<uint32_t> start, cost;
  std::vector<uint32_t> len_at;  // counting sort of a region's reads by length
};
PackScratch& pack_scratch() {
  thread_local PackScratch s;
  return s;
}
}  // namespace

// Appends tiles to dst ordered by cost (anti-diagonal steps of the tile),
// largest first, so that the persistent waves end on small tiles. Costs are
// bucketed by 16 steps; the order inside a bucket is kept.
static void append_largest_first(const std::vector<Tile>& tiles, const std::vector<uint32_t>& cost,
                                 std::vector<Tile>& dst) {
  if (tiles.empty()) return;
  const size_t base = dst.size();
  dst.resize(base + tiles.size());
  // already in order (one region's tiles often all cost the same): a copy
  bool ordered = true;
  for (size_t x = 1; ordered && x < tiles.size(); ++x) ordered = (cost[x - 1] >> 4) >= (cost[x] >> 4);
  if (ordered) {
    std::copy(tiles.begin(), tiles.end(), dst.begin() + base);
    return;
  }
  if (tiles.size() < 512) {
    std::vector<uint32_t>& idx = pack_scratch().idx;
    idx.resize(tiles.size());
    for (uint32_t x = 0; x < idx.size(); ++x) idx[x] = x;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t p, uint32_t q) { return (cost[p] >> 4) > (cost[q] >> 4); });
    for (size_t x = 0; x < idx.size(); ++x) dst[base + x] = tiles[idx[x]];
    return;
  }
  constexpr uint32_t kBuckets = 8192;
  std::vector<uint32_t> count(kBuckets + 1, 0);
  auto bucket = [&](size_t x) { return kBuckets - std::min<uint32_t>(cost[x] >> 4, kBuckets); };
  for (size_t x = 0; x < tiles.size(); ++x) count[bucket(x)]++;
  uint32_t acc = 0;
  for (auto& c : count) {
    const uint32_t n = c;
    c = acc;
    acc += n;
  }
  for (size_t x = 0; x < tiles.size(); ++x) dst[base + count[bucket(x)]++] = tiles[x];
}

void HostPack::reset() {
  HostPack fresh;  // the scalars of a new pack
  auto keep = [](auto& v, auto& into) {
    v.clear();
    into.swap(v);
  };
  keep(region, fresh.region);
  keep(hap_pack, fresh.hap_pack);
  keep(hap_nmask, fresh.hap_nmask);
  keep(hap_off, fresh.hap_off);
  keep(hap_len, fresh.hap_len);
  keep(read_pack, fresh.read_pack);
  keep(read_nmask, fresh.read_nmask);
  keep(read_off, fresh.read_off);
  keep(read_len, fresh.read_len);
  keep(rq, fresh.rq);
  keep(ri, fresh.ri);
  keep(rd, fresh.rd);
  keep(rc, fresh.rc);
  keep(read_perm, fresh.read_perm);
  keep(tiles, fresh.tiles);
  keep(bin_start, fresh.bin_start);
  keep(fp64_cap, fresh.fp64_cap);
  keep(fp64_base, fresh.fp64_base);
  keep(read_raw, fresh.read_raw);
  keep(read_prep, fresh.read_prep);
  keep(staging, fresh.staging);
  keep(nan_scratch, fresh.nan_scratch);
  keep(flag_scratch, fresh.flag_scratch);
  *this = std::move(fresh);
}

int build_host_pack(const rvc_region* regions, uint32_t n_regions, const FP32Policy& pol, HostPack* hp,
                    const DevicePrep* prep) {
  HostPack& P = *hp;
  P.reset();
  PackScratch& W = pack_scratch();
  P.n_regions = n_regions;
  P.region.resize(n_regions);
  uint64_t n_hap_total = 0, n_read_total = 0, hap_bases = 0, read_bases = 0, pairs = 0;
  for (uint32_t r = 0; r < n_regions; ++r) {
    const rvc_region& g = regions[r];
    RegionLayout& L = P.region[r];
    L.n_haps = g.n_haps;
    L.n_reads = g.n_reads;
    L.hap0 = (uint32_t)n_hap_total;
    L.read0 = (uint32_t)n_read_total;
    L.pair0 = pairs;
    n_hap_total += g.n_haps;
    n_read_total += g.n_reads;
    pairs += (uint64_t)g.n_haps * g.n_reads;
    for (uint32_t h = 0; h < g.n_haps; ++h) hap_bases += align_up((uint32_t)(g.hap_off[h + 1] - g.hap_off[h]), kSeqAlign);
    for (uint32_t x = 0; x < g.n_reads; ++x)
      read_bases += align_up((uint32_t)(g.read_off[x + 1] - g.read_off[x]), kSeqAlign);
  }
  if (hap_bases > 0xffffffffull - kSeqAlign || read_bases > 0xffffffffull - kSeqAlign || pairs > 0xffffffffull)
    return set_error(RVC_E_LIMIT, "batch too large for 32-bit offsets");
  P.n_pairs_slots = pairs;
  P.hap_off.resize(n_hap_total);
  P.hap_len.resize(n_hap_total);
  P.read_off.resize(n_read_total);
  P.read_len.resize(n_read_total);
  P.read_perm.resize(n_read_total);
  const uint32_t hw = (uint32_t)(hap_bases / 16 + 2), hn = (uint32_t)(hap_bases / 32 + 2);
  const uint32_t rw = (uint32_t)(read_bases / 16 + 2), rn = (uint32_t)(read_bases / 32 + 2);
  P.hap_pack.assign(hw, 0);
  P.hap_nmask.assign(hn, 0);
  P.read_pack.assign(rw, 0);
  P.read_nmask.assign(rn, 0);
  P.rq.assign(read_bases + 64, 0);
  P.ri.assign(read_bases + 64, 0);
  P.rd.assign(read_bases + 64, 0);
  P.rc.assign(read_bases + 64, 0);
  P.read_prep.assign(n_read_total, 0);
  if (prep) {
    P.any_prep = true;
    P.read_raw.assign(read_bases + 64, 0);
  }
  uint32_t hcur = 0, rcur = 0;
  P.fp64_cap.assign(kNumFp64Lists, 0);
  P.waves_per_cu = pol.waves_per_cu;
  P.prior_table = pol.prior_table;
  P.hap_run = std::max<uint32_t>(1, std::min<uint32_t>(kMaxHapRun, pol.hap_run));
  P.hap_run_fixed = pol.hap_run_fixed;
  P.no_fold = !pol.fold_small_bins;
  for (int b = 0; b < kFp32Bins; ++b) {
    int W, RPL;
    fp32_variant_shape(b % kFp32Variants, &W, &RPL);
    P.occ[b] = (uint8_t)pol.occupancy(W, RPL, bin_class(b) == kClassTable);
  }
  std::vector<std::vector<Tile>>& bins = W.bins;
  std::vector<std::vector<uint32_t>>& bin_cost = W.bin_cost;
  for (int v = 0; v < kFp32Bins; ++v) {
    bins[v].clear();
    bin_cost[v].clear();
  }
  const uint32_t run_max = P.hap_run;
  std::vector<uint8_t>& hap_has_n = W.hap_has_n;
  for (uint32_t r = 0; r < n_regions; ++r) {
    const rvc_region& g = regions[r];
    RegionLayout& L = P.region[r];
    hap_has_n.assign(g.n_haps, 0);
    for (uint32_t h = 0; h < g.n_haps; ++h) {
      const uint32_t len = (uint32_t)(g.hap_off[h + 1] - g.hap_off[h]);
      P.hap_off[L.hap0 + h] = hcur;
      P.hap_len[L.hap0 + h] = len;
      pack_seq(g.hap_bases + g.hap_off[h], len, hcur, P.hap_pack, P.hap_nmask);
      for (uint32_t w = hcur >> 5; w < (hcur + len + 31) >> 5; ++w)
        if (P.hap_nmask[w]) hap_has_n[h] = 1;
      hcur += align_up(len, kSeqAlign);
    }
    // reads of a region in descending length, so that tiles group equal
    // lengths, in input order among equal lengths: as they come when they
    // already descend, else by a counting sort over the lengths (at most
    // kMaxReadLen in a validated region; a longer one falls back to a sort)
    std::vector<uint32_t>& order = W.order;
    order.resize(g.n_reads);
    auto rlen = [&](uint32_t x) { return (uint32_t)(g.read_off[x + 1] - g.read_off[x]); };
    bool descending = true;
    uint32_t longest = 0;
    for (uint32_t x = 0; x < g.n_reads; ++x) {
      longest = std::max(longest, rlen(x));
      if (x > 0 && rlen(x) > rlen(x - 1)) descending = false;
    }
    for (uint32_t x = 0; x < g.n_reads; ++x) order[x] = x;
    if (!descending && longest <= (uint32_t)kMaxReadLen) {
      std::vector<uint32_t>& at = W.len_at;
      at.assign(kMaxReadLen + 2, 0);
      for (uint32_t x = 0; x < g.n_reads; ++x) ++at[kMaxReadLen - rlen(x) + 1];
      for (int l = 1; l <= kMaxReadLen + 1; ++l) at[l] += at[l - 1];
      for (uint32_t x = 0; x < g.n_reads; ++x) order[at[kMaxReadLen - rlen(x)]++] = x;
    } else if (!descending) {
      std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return rlen(a) > rlen(b); });
    }
    // range of the region's gap-open qualities (insertion and deletion) and
    // its smallest gap-continuation quality: they fix the region's form
    uint8_t gop_lo = 127, gop_hi = 0, gcp_lo = 127;
    for (uint32_t sx = 0; sx < g.n_reads; ++sx) {
      const uint32_t x = order[sx];
      const uint32_t len = (uint32_t)(g.read_off[x + 1] - g.read_off[x]);
      const uint32_t gi = L.read0 + sx;
      P.read_perm[gi] = x;
      P.read_off[gi] = rcur;
      P.read_len[gi] = len;
      const int32_t o = g.read_off[x];
      pack_seq(g.read_bases + o, len, rcur, P.read_pack, P.read_nmask);
      std::memcpy(&P.rq[rcur], g.q + o, len);
      if (prep) {
        // raw qualities; the device fills q, i, d, c (gap continuation 10 > 0)
        std::memcpy(&P.read_raw[rcur], g.read_bases + o, len);
        P.read_prep[gi] = kPrepOn | ((uint32_t)(prep->pcr & 3) << 16) | ((uint32_t)(prep->min_q & 255) << 8) |
                          (uint32_t)prep->mq[x];
        // gap open 45, or down to 10 under a PCR model; continuation 10
        gop_lo = std::min<uint8_t>(gop_lo, prep->pcr != 0 ? 10 : 45);
        gop_hi = std::max<uint8_t>(gop_hi, 45);
        gcp_lo = std::min<uint8_t>(gcp_lo, 10);
      } else {
        std::memcpy(&P.ri[rcur], g.i + o, len);
        std::memcpy(&P.rd[rcur], g.d + o, len);
        std::memcpy(&P.rc[rcur], g.c + o, len);
        gap_ranges(g.i + o, g.d + o, g.c + o, len, &gop_lo, &gop_hi, &gcp_lo);
      }
      rcur += align_up(len, kSeqAlign);
    }
    // tiles: one haplotype x up to 64/W reads of similar length, for the
    // haplotypes without N (prior-table kernel) and then those with N
    // (compare kernel), whose shapes may differ; all in the plain class where
    // the region's gap qualities rule the scaled form out
    const int X = scale_exponent(gop_lo, gop_hi, gcp_lo, pol.fold);
    const bool plain = X == kPlainExponent;
    bool kinds[2] = {false, false};
    for (uint32_t h = 0; h < g.n_haps; ++h)
      if (P.hap_len[L.hap0 + h] > 0) kinds[hap_has_n[h]] = true;
    for (uint8_t kind = 0; kind < 2; ++kind) {
      if (!kinds[kind]) continue;
      const int cls = plain ? kClassPlain : kind ? kClassN : kClassTable;
      uint32_t sx = 0;
      while (sx < g.n_reads) {
        const uint32_t R0 = P.read_len[L.read0 + sx];
        if (R0 == 0) break;  // the rest are empty reads
        int W, RPL;
        pol.pick(R0, &W, &RPL, cls != kClassTable);
        const uint32_t nseg_max = 64u / (uint32_t)W;
        uint32_t nseg = 0;
        while (nseg < nseg_max && sx + nseg < g.n_reads && P.read_len[L.read0 + sx + nseg] > 0) ++nseg;
        const int v0 = fp32_variant_index(W, RPL);
        // the read group's bases and its reads per FP64 bin: a tile's cells
        // are (sum of read lengths) x (sum of haplotype lengths)
        uint64_t rsum = 0;
        uint32_t per_bin[kNumFp64Bins] = {};
        for (uint32_t s = 0; s < nseg; ++s) {
          const uint32_t rl = P.read_len[L.read0 + sx + s];
          rsum += rl;
          ++per_bin[fp64_bin(rl)];
        }
        // runs of consecutive non-empty haplotypes of this kind, at most
        // run_max per tile
        uint32_t h = 0;
        while (h < g.n_haps) {
          if (P.hap_len[L.hap0 + h] == 0 || hap_has_n[h] != kind) {
            ++h;
            continue;
          }
          uint32_t nh = 0, cost = 0;
          while (nh < run_max && h + nh < g.n_haps && P.hap_len[L.hap0 + h + nh] > 0 && hap_has_n[h + nh] == kind) {
            cost += P.hap_len[L.hap0 + h + nh] + (uint32_t)W - 1;
            ++nh;
          }
          const int v = v0 + cls * kFp32Variants;
          Tile t;
          t.hap = L.hap0 + h;
          t.read0 = L.read0 + sx;
          t.out0 = (uint32_t)(L.pair0 + (uint64_t)sx * g.n_haps + h);
          t.meta = tile_meta(nseg, g.n_haps, nh, (uint32_t)X);
          bins[v].push_back(t);
          bin_cost[v].push_back(cost);
          P.bin_units[v] += nh;
          uint64_t hsum = 0;
          for (uint32_t hh = h; hh < h + nh; ++hh) hsum += P.hap_len[L.hap0 + hh];
          for (int b = 0; b < kNumFp64Bins; ++b) P.fp64_cap[cls * kNumFp64Bins + b] += per_bin[b] * nh;
          P.n_pairs += (uint64_t)nseg * nh;
          P.n_cells += rsum * hsum;
          P.variant_cells[v] += rsum * hsum;
          h += nh;
        }
        sx += nseg;
      }
    }
  }
  P.tiles.clear();
  P.bin_start.assign(kFp32Bins + 1, 0);
  for (int v = 0; v < kFp32Bins; ++v) {
    P.bin_start[v] = (uint32_t)P.tiles.size();
    append_largest_first(bins[v], bin_cost[v], P.tiles);
  }
  P.bin_start[kFp32Bins] = (uint32_t)P.tiles.size();
  P.fp64_base.assign(kNumFp64Lists + 1, 0);
  for (int b = 0; b < kNumFp64Lists; ++b) P.fp64_base[b + 1] = P.fp64_base[b] + P.fp64_cap[b];
  return RVC_OK;
}

int scale_exponent(int gop_lo, int gop_hi, int gcp_lo, bool allow_scaled) {
  // Gap open and continuation >= 6 keep tMM >= 1/2 and tDD <= 1/4. I~ = I/tMI
  // exceeds M by at most about the largest over the smallest gap-open
  // probability, 10^(spread/10) (at most 2^16 for a spread of 48): X lowers
  // 2^120 by that many bits plus 2, so that no cell holds more than in the
  // 2^120 scale of the plain form.
  if (!allow_scaled) return kPlainExponent;
  if (gop_hi < gop_lo) return 118;  // no read rows
  const int spread = gop_hi - gop_lo;
  if (gop_lo < 6 || gcp_lo < 6 || spread > 48) return kPlainExponent;
  return 118 - (int)std::ceil(spread * 0.33219280948873623);
}

uint32_t fp32_grid(const HostPack& P, int v) {
  int W, RPL;
  fp32_variant_shape(v % kFp32Variants, &W, &RPL);
  const bool sc = bin_class(v) != kClassPlain;
  const bool pt = bin_class(v) == kClassTable && P.prior_table;
  const int occ = pt ? P.occ[v] : P.occ[v % kFp32Variants];
  int per_cu = std::max(1, fp32_blocks_per_cu(W, RPL, sc, occ, pt));
  if (P.waves_per_cu) per_cu = std::min<int>(per_cu, (int)P.waves_per_cu);
  return (uint32_t)per_cu * (uint32_t)g_engine.n_cu;
}

uint32_t run_for_bin(const HostPack& P, int v, uint64_t units) {
  // about 12 tiles per persistent wave: measured on MI355X against fixed
  // runs, 400 regions ran 5 % faster than at 32 tiles per wave (runs of 4
  // haplotypes instead of 1); batches of 4096 regions keep the longest runs
  constexpr uint64_t kTilesPerWave = 12;
  if (P.hap_run <= 1 || units == 0) return 1;
  if (P.hap_run_fixed) return P.hap_run;
  const uint64_t run = units / (kTilesPerWave * fp32_grid(P, v));
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(P.hap_run, run));
}

namespace {
// Appends the pieces of tile t with at most 'run' haplotypes each.
template <typename F>
void split_tile(const Tile& t, uint32_t run, F&& emit) {
  const uint32_t nh = tile_nh(t.meta);
  for (uint32_t h = 0; h < nh; h += run) {
    Tile p = t;
    p.hap = t.hap + h;
    p.out0 = t.out0 + h;
    p.meta = tile_with_nh(t.meta, std::min(run, nh - h));
    emit(p);
  }
}
}  // namespace

void fold_bins(HostPack& P, int (&to)[kFp32Bins]) {
  // A bin of fewer cells than this runs well below the large bins' rate (on
  // MI355X: bins of 2e9-1.4e10 cells at 5.0-6.2 k GCUPS, bins of 2e11+ at
  // 9 k): its launch's ramp and drain, and runs of one haplotype per tile.
  // The next larger shape of its class and width sweeps its tiles with at
  // most a third more rows (padding rows, which leave the results as they
  // are), at least as fast.
  const uint64_t small = 80000000ull * (uint64_t)std::max(1, g_engine.n_cu);
  for (int v = 0; v < kFp32Bins; ++v) to[v] = v;
  for (int cls = 0; cls < kNumClasses; ++cls)
    for (int w = 0; w < 4; ++w) {
      int big = -1, big_rpl = 0;
      for (int rpl = kMaxRpl; rpl >= 1; --rpl) {
        const int v = cls * kFp32Variants + w * kMaxRpl + rpl - 1;
        if (P.bin_units[v] == 0) continue;
        if (big >= 0 && 3 * big_rpl <= 4 * rpl && P.variant_cells[v] < small) {
          to[v] = big;
          continue;
        }
        big = v;
        big_rpl = rpl;
      }
    }
  if (P.no_fold)
    for (int v = 0; v < kFp32Bins; ++v) to[v] = v;
  for (int v = 0; v < kFp32Bins; ++v)
    if (to[v] != v) {
      P.variant_cells[to[v]] += P.variant_cells[v];
      P.bin_units[to[v]] += P.bin_units[v];
      P.variant_cells[v] = 0;
      P.bin_units[v] = 0;
    }
}

void split_runs(HostPack& P) {
  PackScratch& S = pack_scratch();
  std::vector<Tile>& out = S.out;
  std::vector<uint32_t>& start = S.start;
  std::vector<Tile>& bin = S.bin;
  std::vector<uint32_t>& cost = S.cost;
  int to[kFp32Bins];
  fold_bins(P, to);
  // the bins holding tiles, grouped by the bin they are swept in (a counting
  // sort, increasing within a group): a one-region call touches a handful of
  // the bins, so nothing here may cost bins x bins
  int first[kFp32Bins + 1] = {};
  int src[kFp32Bins];
  for (int u = 0; u < kFp32Bins; ++u)
    if (P.bin_start[u] < P.bin_start[u + 1]) ++first[to[u] + 1];
  for (int v = 0; v < kFp32Bins; ++v) first[v + 1] += first[v];
  {
    int fill[kFp32Bins];
    std::memcpy(fill, first, sizeof fill);
    for (int u = 0; u < kFp32Bins; ++u)
      if (P.bin_start[u] < P.bin_start[u + 1]) src[fill[to[u]]++] = u;
  }
  out.clear();
  start.assign(kFp32Bins + 1, 0);
  for (int v = 0; v < kFp32Bins; ++v) {
    start[v] = (uint32_t)out.size();
    if (first[v] == first[v + 1]) continue;
    const uint32_t run = run_for_bin(P, v, P.bin_units[v]);
    int W, RPL;
    fp32_variant_shape(v % kFp32Variants, &W, &RPL);
    bin.clear();
    cost.clear();
    // this bin's own tiles and those of the bins folded into it
    for (int i = first[v]; i < first[v + 1]; ++i) {
      const int u = src[i];
      for (uint32_t x = P.bin_start[u]; x < P.bin_start[u + 1]; ++x)
        split_tile(P.tiles[x], run, [&](const Tile& p) {
          uint32_t c = 0;
          for (uint32_t h = 0; h < tile_nh(p.meta); ++h) c += P.hap_len[p.hap + h] + (uint32_t)W - 1;
          bin.push_back(p);
          cost.push_back(c);
        });
    }
    append_largest_first(bin, cost, out);
  }
  start[kFp32Bins] = (uint32_t)out.size();
  P.tiles.swap(out);
  P.bin_start.swap(start);
}

int fp32_variant_index(int W, int RPL) {
  const int w = W == 16 ? 0 : W == 32 ? 1 : W == 64 ? 2 : 3;
  return w * kMaxRpl + (RPL - 1);
}
void fp32_variant_shape(int v, int* W, int* RPL) {
  static const int kW[4] = {16, 32, 64, 8};
  *W = kW[v / kMaxRpl];
  *RPL = v % kMaxRpl + 1;
}

// ---------------------------------------------------------------------------
// device batch
// ---------------------------------------------------------------------------
namespace {
template <typename T>
size_t bytes_of(const std::vector<T>& v) { return v.size() * sizeof(T); }
inline size_t round256(size_t x) { return (x + 255) & ~size_t(255); }
}  // namespace

DeviceBatch::~DeviceBatch() { release(); }

void DeviceBatch::release() {
  if (graph_exec) hipGraphExecDestroy(graph_exec);
  if (graph) hipGraphDestroy(graph);
  graph_exec = nullptr;
  graph = nullptr;
  for (auto& e : ev) if (e) hipEventDestroy(e);
  for (auto& e : ev) e = nullptr;
  for (auto& e : lev) if (e) hipEventDestroy(e);
  lev.clear();
  if (own_stream && stream) hipStreamDestroy(stream);
  stream = nullptr;
  if (mem && own_mem) hipFree(mem);
  mem = nullptr;
}

PackSizes pack_sizes(const HostPack& P) {
  PackSizes z;
  z.hap_words = P.hap_pack.size();
  z.hap_nwords = P.hap_nmask.size();
  z.n_haps = P.hap_off.size();
  z.read_words = P.read_pack.size();
  z.read_nwords = P.read_nmask.size();
  z.qual_bytes = P.rq.size();
  z.n_reads = P.read_off.size();
  z.n_tiles = P.tiles.size();
  z.n_pairs_slots = P.n_pairs_slots;
  z.n_rescue = P.fp64_base.empty() ? 0 : P.fp64_base[kNumFp64Lists];
  z.raw_bytes = P.any_prep ? P.read_raw.size() : 0;
  return z;
}

BatchLayout batch_layout(const PackSizes& z) {
  BatchLayout L;
  size_t off = 0;
  auto take = [&](size_t b) { size_t o = off; off += round256(b); return o; };
  L.hpack = take(4 * z.hap_words);
  L.hn = take(4 * z.hap_nwords);
  L.hoff = take(4 * z.n_haps);
  L.hlen = take(4 * z.n_haps);
  L.rpack = take(4 * z.read_words);
  L.rn = take(4 * z.read_nwords);
  L.rq = take(z.qual_bytes);
  L.ri = take(z.qual_bytes);
  L.rd = take(z.qual_bytes);
  L.rc = take(z.qual_bytes);
  L.roff = take(4 * z.n_reads);
  L.rlen = take(4 * z.n_reads);
  L.tiles = take(sizeof(Tile) * (z.n_tiles + 1));
  L.rbase = take(4 * kNumFp64Lists);
  L.raw = take(z.raw_bytes);
  L.prep = take(4 * z.n_reads);
  L.in_bytes = off;
  // flags right behind the results: one copy brings both back
  // results, flags (rounded up to 8 bytes) and the counters one after
  // another: one copy brings all three back
  const size_t flag_bytes = (z.n_pairs_slots + 7) & ~size_t(7);
  L.out = take(sizeof(double) * z.n_pairs_slots + flag_bytes + 4 * kNumCounters);
  L.flags = L.out + sizeof(double) * z.n_pairs_slots;
  L.cnt = L.flags + flag_bytes;
  L.resc = take(sizeof(RescueEntry) * (z.n_rescue + 1));
  L.bytes = off;
  return L;
}

void stage_pack(const HostPack& P, const BatchLayout& L, char* dst) {
  auto put = [&](size_t o, const void* src, size_t b) { if (b) std::memcpy(dst + o, src, b); };
  put(L.hpack, P.hap_pack.data(), bytes_of(P.hap_pack));
  put(L.hn, P.hap_nmask.data(), bytes_of(P.hap_nmask));
  put(L.hoff, P.hap_off.data(), bytes_of(P.hap_off));
  put(L.hlen, P.hap_len.data(), bytes_of(P.hap_len));
  put(L.rpack, P.read_pack.data(), bytes_of(P.read_pack));
  put(L.rn, P.read_nmask.data(), bytes_of(P.read_nmask));
  put(L.rq, P.rq.data(), P.rq.size());
  put(L.ri, P.ri.data(), P.ri.size());
  put(L.rd, P.rd.data(), P.rd.size());
  put(L.rc, P.rc.data(), P.rc.size());
  put(L.roff, P.read_off.data(), bytes_of(P.read_off));
  put(L.rlen, P.read_len.data(), bytes_of(P.read_len));
  put(L.tiles, P.tiles.data(), bytes_of(P.tiles));
  put(L.rbase, P.fp64_base.data(), kNumFp64Lists * sizeof(uint32_t));
  if (P.any_prep) put(L.raw, P.read_raw.data(), P.read_raw.size());
  put(L.prep, P.read_prep.data(), bytes_of(P.read_prep));
}

size_t stage_results(const HostPack& P, const BatchLayout& L, char* dst) {
  const size_t n = P.n_pairs_slots;
  const size_t flag_bytes = (n + 7) & ~size_t(7);
  const uint64_t nan = 0x7ff8000000000000ull;
  uint64_t* o = reinterpret_cast<uint64_t*>(dst + L.out);
  for (size_t x = 0; x < n; ++x) o[x] = nan;
  std::memset(dst + L.flags, 0, flag_bytes + 4 * kNumCounters);
  return L.cnt + 4 * kNumCounters;
}

PinnedBuf::~PinnedBuf() { release(); }

void PinnedBuf::release() {
  if (ptr) hipHostFree(ptr);
  for (char* p : retired) hipHostFree(p);
  retired.clear();
  ptr = nullptr;
  cap = 0;
}

int PinnedBuf::reserve(size_t bytes) {
  if (bytes <= cap) return RVC_OK;
  const size_t want = std::max(bytes + bytes / 4 + (1u << 20), 2 * cap);
  if (ptr) retired.push_back(ptr);
  ptr = nullptr;
  cap = 0;
  if (hipHostMalloc((void**)&ptr, want, hipHostMallocDefault) != hipSuccess) {
    ptr = nullptr;
    return set_error(RVC_E_DEVICE, "pinned host buffer of %zu bytes", want);
  }
  cap = want;
  return RVC_OK;
}

int stage_merged(const HostPack* const* packs, uint32_t n, HostPack* meta, BatchLayout* Lout, PinnedBuf* buf) {
  HostPack& P = *meta;
  P = HostPack();
  // every sequence block starts on a 32-base boundary: a pack's words are
  // taken without its two spare words, rounded down to an even count, so the
  // next block still starts on a 32-base boundary
  auto hwords = [](const HostPack& A) { return (A.hap_pack.size() - 2) & ~size_t(1); };
  auto rwords = [](const HostPack& A) { return (A.read_pack.size() - 2) & ~size_t(1); };
  PackSizes z;
  P.fp64_cap.assign(kNumFp64Lists, 0);
  for (uint32_t x = 0; x < n; ++x) {
    const HostPack& A = *packs[x];
    z.hap_words += hwords(A);
    z.n_haps += A.hap_off.size();
    z.read_words += rwords(A);
    z.n_reads += A.read_off.size();
    z.n_pairs_slots += A.n_pairs_slots;
    for (int b = 0; b < kNumFp64Lists; ++b) P.fp64_cap[b] += A.fp64_cap[b];
    for (int v = 0; v < kFp32Bins; ++v) {
      P.bin_units[v] += A.bin_units[v];
      P.variant_cells[v] += A.variant_cells[v];
    }
    P.n_pairs += A.n_pairs;
    P.n_cells += A.n_cells;
    P.waves_per_cu = A.waves_per_cu;
    P.prior_table = A.prior_table;
    P.hap_run = A.hap_run;
    P.hap_run_fixed = A.hap_run_fixed;
    P.any_prep = P.any_prep || A.any_prep;
    std::memcpy(P.occ, A.occ, sizeof P.occ);
  }
  // tiles of the merged batch: small bins folded into larger shapes, each
  // pack's runs split for the batch's size
  P.no_fold = n > 0 && packs[0]->no_fold;
  int to[kFp32Bins];
  fold_bins(P, to);
  uint32_t run[kFp32Bins];
  std::vector<uint64_t> bin_count(kFp32Bins, 0);
  for (int v = 0; v < kFp32Bins; ++v) run[v] = run_for_bin(P, to[v], P.bin_units[to[v]]);
  for (uint32_t x = 0; x < n; ++x) {
    const HostPack& A = *packs[x];
    for (int v = 0; v < kFp32Bins; ++v)
      for (uint32_t k = A.bin_start[v]; k < A.bin_start[v + 1]; ++k)
        bin_count[to[v]] += (tile_nh(A.tiles[k].meta) + run[v] - 1) / run[v];
  }
  for (int v = 0; v < kFp32Bins; ++v) z.n_tiles += bin_count[v];
  if (z.hap_words * 16 > 0xffffffffull || z.read_words * 16 > 0xffffffffull || z.n_pairs_slots > 0xffffffffull)
    return set_error(RVC_E_LIMIT, "merged batch too large for 32-bit offsets");
  z.hap_nwords = z.hap_words / 2 + 2;
  z.hap_words += 2;
  z.qual_bytes = z.read_words * 16 + 64;
  z.read_nwords = z.read_words / 2 + 2;
  z.read_words += 2;
  z.raw_bytes = P.any_prep ? z.qual_bytes : 0;
  P.fp64_base.assign(kNumFp64Lists + 1, 0);
  for (int b = 0; b < kNumFp64Lists; ++b) P.fp64_base[b + 1] = P.fp64_base[b] + P.fp64_cap[b];
  z.n_rescue = P.fp64_base[kNumFp64Lists];
  const BatchLayout L = batch_layout(z);
  int rc = buf->reserve(L.in_bytes);
  if (rc) return rc;
  char* dst = buf->ptr;
  uint32_t* hpack = reinterpret_cast<uint32_t*>(dst + L.hpack);
  uint32_t* hn = reinterpret_cast<uint32_t*>(dst + L.hn);
  uint32_t* hoff = reinterpret_cast<uint32_t*>(dst + L.hoff);
  uint32_t* hlen = reinterpret_cast<uint32_t*>(dst + L.hlen);
  uint32_t* rpack = reinterpret_cast<uint32_t*>(dst + L.rpack);
  uint32_t* rn = reinterpret_cast<uint32_t*>(dst + L.rn);
  uint32_t* roff = reinterpret_cast<uint32_t*>(dst + L.roff);
  uint32_t* rlen = reinterpret_cast<uint32_t*>(dst + L.rlen);
  Tile* tiles = reinterpret_cast<Tile*>(dst + L.tiles);
  char* quals[4] = {dst + L.rq, dst + L.ri, dst + L.rd, dst + L.rc};
  uint8_t* raw = reinterpret_cast<uint8_t*>(dst + L.raw);
  uint32_t* prep = reinterpret_cast<uint32_t*>(dst + L.prep);
  P.bin_start.assign(kFp32Bins + 1, 0);
  for (int v = 0; v < kFp32Bins; ++v) P.bin_start[v + 1] = P.bin_start[v] + (uint32_t)bin_count[v];
  std::vector<uint32_t> bin_cur(P.bin_start.begin(), P.bin_start.end() - 1);
  P.region.reserve(n);
  P.read_perm.resize(z.n_reads);
  uint64_t hw = 0, rw = 0, hidx = 0, ridx = 0, pbase = 0;
  for (uint32_t x = 0; x < n; ++x) {
    const HostPack& A = *packs[x];
    const size_t aw = hwords(A), bw = rwords(A);
    std::memcpy(hpack + hw, A.hap_pack.data(), aw * 4);
    std::memcpy(hn + hw / 2, A.hap_nmask.data(), aw / 2 * 4);
    std::memcpy(rpack + rw, A.read_pack.data(), bw * 4);
    std::memcpy(rn + rw / 2, A.read_nmask.data(), bw / 2 * 4);
    const std::vector<uint8_t>* aq[4] = {&A.rq, &A.ri, &A.rd, &A.rc};
    for (int k = 0; k < 4; ++k) std::memcpy(quals[k] + rw * 16, aq[k]->data(), bw * 16);
    // reads preprocessed on the device: raw bases and their prep words (zero
    // words for reads that arrive preprocessed; their raw bytes are not read)
    if (A.any_prep) std::memcpy(raw + rw * 16, A.read_raw.data(), bw * 16);
    std::memcpy(prep + ridx, A.read_prep.data(), A.read_prep.size() * 4);
    const uint32_t hbase = (uint32_t)(hw * 16), rbase = (uint32_t)(rw * 16);
    for (size_t h = 0; h < A.hap_off.size(); ++h) {
      hoff[hidx + h] = A.hap_off[h] + hbase;
      hlen[hidx + h] = A.hap_len[h];
    }
    for (size_t r = 0; r < A.read_off.size(); ++r) {
      roff[ridx + r] = A.read_off[r] + rbase;
      rlen[ridx + r] = A.read_len[r];
    }
    std::memcpy(P.read_perm.data() + ridx, A.read_perm.data(), A.read_perm.size() * 4);
    for (const RegionLayout& L0 : A.region) {
      RegionLayout R = L0;
      R.hap0 += (uint32_t)hidx;
      R.read0 += (uint32_t)ridx;
      R.pair0 += pbase;
      P.region.push_back(R);
    }
    for (int v = 0; v < kFp32Bins; ++v) {
      Tile* t = tiles + bin_cur[to[v]];
      for (uint32_t k = A.bin_start[v]; k < A.bin_start[v + 1]; ++k) {
        Tile b = A.tiles[k];
        b.hap += (uint32_t)hidx;
        b.read0 += (uint32_t)ridx;
        b.out0 += (uint32_t)pbase;
        split_tile(b, run[v], [&](const Tile& q) { *t++ = q; });
      }
      bin_cur[to[v]] = (uint32_t)(t - tiles);
    }
    hw += aw;
    rw += bw;
    hidx += A.hap_off.size();
    ridx += A.read_off.size();
    pbase += A.n_pairs_slots;
  }
  // spare words and bytes at the ends, read by the last steps of a sweep
  std::memset(hpack + hw, 0, 8);
  std::memset(hn + hw / 2, 0, 8);
  std::memset(rpack + rw, 0, 8);
  std::memset(rn + rw / 2, 0, 8);
  for (int k = 0; k < 4; ++k) std::memset(quals[k] + rw * 16, 0, 64);
  std::memset(static_cast<void*>(tiles + z.n_tiles), 0, sizeof(Tile));
  std::memcpy(dst + L.rbase, P.fp64_base.data(), kNumFp64Lists * sizeof(uint32_t));
  P.n_pairs_slots = z.n_pairs_slots;
  P.n_regions = (uint32_t)P.region.size();
  *Lout = L;
  return RVC_OK;
}

constexpr size_t kKernelUploadMax = 4u << 20;

int DeviceBatch::upload(const HostPack& P, hipStream_t s, void* reuse_mem, size_t reuse_bytes) {
  const BatchLayout L = batch_layout(pack_sizes(P));
  std::vector<char>& img = const_cast<HostPack&>(P).staging;
  img.resize(L.in_bytes);
  stage_pack(P, L, img.data());
  return upload_staged(P, L, img.data(), s, reuse_mem, reuse_bytes);
}

int DeviceBatch::upload_staged(const HostPack& P, const BatchLayout& L, const void* host, hipStream_t s,
                               void* reuse_mem, size_t reuse_bytes, bool with_results, bool src_pinned) {
  pack = &P;
  layout = L;
  bytes = L.bytes;
  if (reuse_mem && reuse_bytes >= bytes) {
    mem = reuse_mem;
    own_mem = false;
  } else {
    RVC_HIP(hipMalloc(&mem, bytes));
    own_mem = true;
  }
  char* base = static_cast<char*>(mem);
  // with_results: the host image carries the starting results and zero
  // counters too (stage_results); else the first launch sequence's begin
  // kernel sets results to NaN with no flags (pairs with an empty side keep
  // that) and clears the counters
  const size_t up = with_results ? L.cnt + 4 * kNumCounters : L.in_bytes;
  // one region's image: a kernel reads the pinned staging (about 3 us sooner
  // than a copy engine whose end the sweep must wait for); larger images go
  // by DMA
  if (src_pinned && up <= kKernelUploadMax) RVC_HIP(launch_copy(base, host, up, s));
  else RVC_HIP(hipMemcpyAsync(base, host, up, hipMemcpyHostToDevice, s));
  init_pending = !with_results;
  counters_clear = with_results;
  seqs.hap_pack = reinterpret_cast<const uint32_t*>(base + L.hpack);
  seqs.hap_nmask = reinterpret_cast<const uint32_t*>(base + L.hn);
  seqs.hap_off = reinterpret_cast<const uint32_t*>(base + L.hoff);
  seqs.hap_len = reinterpret_cast<const uint32_t*>(base + L.hlen);
  seqs.read_pack = reinterpret_cast<const uint32_t*>(base + L.rpack);
  seqs.read_nmask = reinterpret_cast<const uint32_t*>(base + L.rn);
  seqs.rq = reinterpret_cast<const uint8_t*>(base + L.rq);
  seqs.ri = reinterpret_cast<const uint8_t*>(base + L.ri);
  seqs.rd = reinterpret_cast<const uint8_t*>(base + L.rd);
  seqs.rc = reinterpret_cast<const uint8_t*>(base + L.rc);
  seqs.read_off = reinterpret_cast<const uint32_t*>(base + L.roff);
  seqs.read_len = reinterpret_cast<const uint32_t*>(base + L.rlen);
  tiles = reinterpret_cast<const Tile*>(base + L.tiles);
  out = reinterpret_cast<double*>(base + L.out);
  flags = reinterpret_cast<uint8_t*>(base + L.flags);
  rescue = reinterpret_cast<RescueEntry*>(base + L.resc);
  rescue_base = reinterpret_cast<const uint32_t*>(base + L.rbase);
  counters = reinterpret_cast<uint32_t*>(base + L.cnt);
  n_counters_ = kNumCounters;
  read_raw = reinterpret_cast<const uint8_t*>(base + L.raw);
  read_prep = reinterpret_cast<const uint32_t*>(base + L.prep);
  return RVC_OK;
}

int DeviceBatch::enqueue(hipStream_t s, bool timed, const uint32_t* rescued) {
  const HostPack& P = *pack;
  uint32_t* heads = counters;
  uint32_t* rescue_count = counters + kFp32Bins + kNumFp64Lists;
  // one kernel: counters to zero and, after an upload, results to NaN / flags 0
  // (none where the upload brought both)
  if (init_pending || !counters_clear)
    RVC_HIP(launch_begin(counters, n_counters_, out, init_pending ? P.n_pairs_slots : 0, s));
  init_pending = false;
  counters_clear = false;
  if (P.any_prep) {
    // raw qualities -> q, i, d, c in place (idempotent: a replay gives the same)
    PrepArgs pa;
    pa.rq = const_cast<uint8_t*>(seqs.rq);
    pa.ri = const_cast<uint8_t*>(seqs.ri);
    pa.rd = const_cast<uint8_t*>(seqs.rd);
    pa.rc = const_cast<uint8_t*>(seqs.rc);
    pa.raw = read_raw;
    pa.prep = read_prep;
    pa.read_off = seqs.read_off;
    pa.read_len = seqs.read_len;
    pa.gop_cap = g_engine.tables.gop_cap;
    pa.n_reads = (uint32_t)(P.read_perm.size());
    RVC_HIP(launch_prep(pa, s));
  }
  if (timed) RVC_HIP(hipEventRecord(ev[0], s));
  n_fp32_launches = 0;
  n_fp64_launches = 0;
  launches.clear();
  auto mark = [&](int idx) -> int {
    if (!timed) return RVC_OK;
    while ((int)lev.size() <= idx) {
      hipEvent_t e;
      RVC_HIP(hipEventCreate(&e));
      lev.push_back(e);
    }
    RVC_HIP(hipEventRecord(lev[idx], s));
    return RVC_OK;
  };
  int rc = mark(0);
  if (rc) return rc;
  for (int v = 0; v < kFp32Bins; ++v) {
    const uint32_t n = P.bin_start[v + 1] - P.bin_start[v];
    if (n == 0) continue;
    int W, RPL;
    fp32_variant_shape(v % kFp32Variants, &W, &RPL);
    const bool sc = bin_class(v) != kClassPlain;
    const bool pt = bin_class(v) == kClassTable && P.prior_table;
    Fp32Args a;
    a.s = seqs;
    a.t = g_engine.tables;
    a.tiles = tiles + P.bin_start[v];
    a.n_tiles = n;
    a.head = heads + v;
    a.out = out;
    a.flags = flags;
    a.rescue = rescue;
    a.rescue_count = rescue_count;
    a.rescue_base = rescue_base;
    a.rescue_list0 = (uint32_t)(bin_class(v) * kNumFp64Bins);
    const int occ = pt ? P.occ[v] : P.occ[v % kFp32Variants];
    const int grid = (int)std::min<uint64_t>(n, fp32_grid(P, v));
    hipError_t e = launch_fp32(W, RPL, sc, occ, pt, grid, s, a);
    if (e != hipSuccess) return set_error(RVC_E_DEVICE, "FP32 launch W=%d RPL=%d: %s", W, RPL, hipGetErrorString(e));
    ++n_fp32_launches;
    rvc_kernel_time kt{};
    snprintf(kt.name, sizeof kt.name, "k_fp32%s<%d,%d,%s,occ%d>", pt ? "pt" : "", W, RPL,
             sc ? "scaled" : "plain", occ);
    kt.precision = 32; kt.width = W; kt.rows_per_lane = RPL; kt.occupancy = P.occ[v];
    kt.tiles = n; kt.cells = P.variant_cells[v];
    launches.push_back(kt);
    if ((rc = mark((int)launches.size()))) return rc;
  }
  if (timed) RVC_HIP(hipEventRecord(ev[1], s));
  if (rescued == kNoFp64) {
    if (timed) RVC_HIP(hipEventRecord(ev[2], s));
    return RVC_OK;
  }
  return enqueue_fp64(s, timed, rescued);
}

int DeviceBatch::enqueue_fp64(hipStream_t s, bool timed, const uint32_t* rescued) {
  const HostPack& P = *pack;
  uint32_t* heads = counters;
  uint32_t* rescue_count = counters + kFp32Bins + kNumFp64Lists;
  int rc = RVC_OK;
  auto mark = [&](int idx) -> int {
    if (!timed) return RVC_OK;
    while ((int)lev.size() <= idx) {
      hipEvent_t e;
      RVC_HIP(hipEventCreate(&e));
      lev.push_back(e);
    }
    RVC_HIP(hipEventRecord(lev[idx], s));
    return RVC_OK;
  };
  for (int l = 0; l < kNumFp64Lists; ++l) {
    // every list that can hold pairs, or (with the counts read back after the
    // FP32 pass) only the lists that do
    const uint64_t n_list = rescued ? rescued[l] : P.fp64_cap[l];
    if (n_list == 0) continue;
    const int b = l % kNumFp64Bins;
    const int cls = l / kNumFp64Bins;
    const bool sc = cls != kClassPlain;
    const bool pt = cls == kClassTable && P.prior_table;
    Fp64Args a;
    a.s = seqs;
    a.t = g_engine.tables;
    a.list = rescue + P.fp64_base[l];
    a.count = rescue_count + l;
    a.head = heads + kFp32Bins + l;
    a.out = out;
    a.flags = flags;
    const int per_cu = std::max(1, fp64_blocks_per_cu(b, pt, sc));
    int W, RPL;
    fp64_bin_shape(b, &W, &RPL);
    const uint64_t max_waves = (n_list + (64 / W) - 1) / (64 / W);
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(max_waves, (uint64_t)per_cu * g_engine.n_cu));
    hipError_t e = launch_fp64(b, pt, sc, grid, s, a);
    if (e != hipSuccess) return set_error(RVC_E_DEVICE, "FP64 launch list %d: %s", l, hipGetErrorString(e));
    ++n_fp64_launches;
    rvc_kernel_time kt{};
    snprintf(kt.name, sizeof kt.name, "k_fp64%s<%d,%d,%s>", pt ? "pt" : "", W, RPL, sc ? "scaled" : "plain");
    kt.precision = 64; kt.width = W; kt.rows_per_lane = RPL; kt.occupancy = 0;
    kt.tiles = P.fp64_cap[l];
    launches.push_back(kt);
    if ((rc = mark((int)launches.size()))) return rc;
  }
  if (timed) RVC_HIP(hipEventRecord(ev[2], s));
  return RVC_OK;
}

size_t DeviceBatch::fetch_bytes() const {
  const size_t n = pack->n_pairs_slots;
  return sizeof(double) * n + ((n + 7) & ~size_t(7)) + 4 * kNumCounters;
}

int DeviceBatch::run_and_fetch(hipStream_t s, char* host, bool host_pinned) {
  const size_t n = pack->n_pairs_slots;
  int rc = enqueue(s, false, kNoFp64);
  if (rc) return rc;
  if (host_pinned) RVC_HIP(launch_copy(host, out, fetch_bytes(), s));
  else RVC_HIP(hipMemcpyAsync(host, out, fetch_bytes(), hipMemcpyDeviceToHost, s));
  RVC_HIP(hipStreamSynchronize(s));
  const uint32_t* cnt = reinterpret_cast<const uint32_t*>(host + sizeof(double) * n + ((n + 7) & ~size_t(7)));
  uint32_t rescued[kNumFp64Lists];
  std::memcpy(rescued, cnt + kFp32Bins + kNumFp64Lists, sizeof rescued);
  bool any = false;
  for (uint32_t c : rescued) any = any || c != 0;
  n_rescued = 0;
  for (uint32_t c : rescued) n_rescued += c;
  if (!any) return RVC_OK;
  if ((rc = enqueue_fp64(s, false, rescued))) return rc;
  if (host_pinned) RVC_HIP(launch_copy(host, out, n * (sizeof(double) + 1), s));
  else RVC_HIP(hipMemcpyAsync(host, out, n * (sizeof(double) + 1), hipMemcpyDeviceToHost, s));
  RVC_HIP(hipStreamSynchronize(s));
  return RVC_OK;
}

int DeviceBatch::ensure_events() {
  for (auto& e : ev)
    if (!e) RVC_HIP(hipEventCreate(&e));
  return RVC_OK;
}

int DeviceBatch::run(bool use_graph_) {
  int rc = ensure_events();
  if (rc) return rc;
  if (use_graph_) {
    if (!graph_exec) {
      RVC_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue(stream, false);
      launches.clear();
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(stream, &gr);
      if (rc) return rc;
      if (e != hipSuccess) return set_error(RVC_E_DEVICE, "graph capture: %s", hipGetErrorString(e));
      graph = gr;
      RVC_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    }
    RVC_HIP(hipEventRecord(ev[0], stream));
    RVC_HIP(hipGraphLaunch(graph_exec, stream));
    RVC_HIP(hipEventRecord(ev[2], stream));
    RVC_HIP(hipEventSynchronize(ev[2]));
    float ms = 0;
    RVC_HIP(hipEventElapsedTime(&ms, ev[0], ev[2]));
    total_ms = ms;
    fp32_ms = fp64_ms = -1;
  } else if (untimed) {
    // the same launches as the graph, queued one by one, with no events
    // between them
    RVC_HIP(hipEventRecord(ev[0], stream));
    rc = enqueue(stream, false);
    if (rc) return rc;
    RVC_HIP(hipEventRecord(ev[2], stream));
    RVC_HIP(hipEventSynchronize(ev[2]));
    float ms = 0;
    RVC_HIP(hipEventElapsedTime(&ms, ev[0], ev[2]));
    total_ms = ms;
    fp32_ms = fp64_ms = -1;
  } else {
    rc = enqueue(stream, true);
    if (rc) return rc;
    RVC_HIP(hipEventSynchronize(ev[2]));
    float a = 0, b = 0;
    RVC_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
    RVC_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
    fp32_ms = a;
    fp64_ms = b;
    total_ms = a + b;
    for (size_t x = 0; x < launches.size(); ++x) {
      float m = 0;
      RVC_HIP(hipEventElapsedTime(&m, lev[x], lev[x + 1]));
      launches[x].ms = m;
    }
  }
  // rescued pair count of this run
  uint32_t cnt[kNumFp64Lists];
  RVC_HIP(hipMemcpy(cnt, counters + kFp32Bins + kNumFp64Lists, sizeof cnt, hipMemcpyDeviceToHost));
  n_rescued = 0;
  for (int b = 0; b < kNumFp64Lists; ++b) n_rescued += cnt[b];
  return RVC_OK;
}

int DeviceBatch::download(double* host_out, uint8_t* host_flags, hipStream_t s) {
  const HostPack& P = *pack;
  std::vector<double>& tmp = const_cast<HostPack&>(P).nan_scratch;
  std::vector<uint8_t>& tf = const_cast<HostPack&>(P).flag_scratch;
  tmp.resize(P.n_pairs_slots);
  tf.resize(P.n_pairs_slots);
  RVC_HIP(hipMemcpyAsync(tmp.data(), out, P.n_pairs_slots * sizeof(double), hipMemcpyDeviceToHost, s));
  RVC_HIP(hipMemcpyAsync(tf.data(), flags, P.n_pairs_slots, hipMemcpyDeviceToHost, s));
  RVC_HIP(hipStreamSynchronize(s));
  unpermute_results(P, tmp.data(), tf.data(), host_out, host_flags);
  return RVC_OK;
}

void unpermute_results(const HostPack& P, const double* dev_out, const uint8_t* dev_flags, double* host_out,
                       uint8_t* host_flags) {
  // device rows are in descending read length; put them back in input order
  for (uint32_t r = 0; r < P.n_regions; ++r) {
    const RegionLayout& L = P.region[r];
    for (uint32_t sx = 0; sx < L.n_reads; ++sx) {
      const uint32_t x = P.read_perm[L.read0 + sx];
      const size_t src = L.pair0 + (size_t)sx * L.n_haps, dst = L.pair0 + (size_t)x * L.n_haps;
      if (host_out) std::memcpy(host_out + dst, dev_out + src, L.n_haps * sizeof(double));
      if (host_flags) std::memcpy(host_flags + dst, dev_flags + src, L.n_haps);
    }
  }
}

}  // namespace rvc
