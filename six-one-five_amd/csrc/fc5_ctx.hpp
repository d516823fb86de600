// fc5_ctx.hpp -- internal state behind the C ABI (contexts, bands, device
// tables) shared by the ABI translation units.  Not a public header.
#pragma once

#include "fc5.h"

#include <hip/hip_runtime.h>

#include <functional>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "fc5_host.hpp"
#include "fc5_kernels.hpp"
#include "fc5_twin.hpp"

namespace fc5 {

struct PartBufs;  // partitioned probing buffers (fc5_part_api.cpp)

constexpr u64 kGateModU = 124488;
constexpr u32 kGateWordsU = (124488 + 31) / 32;

struct DeviceTable {
  StoreHeader hdr{};
  u64* cells = nullptr;
  XorDev dev{};
  bool present = false;
};

inline void free_table(DeviceTable& t) {
  if (t.cells) (void)hipFree(t.cells);
  t = DeviceTable{};
}

// Cell-gather cache policy of a table: the option when set, else plain loads
// for tables of up to 1 GiB of cells (they partly stay in L2 and the MALL) and
// non-temporal loads above (lines are not reused; measured 1-12% faster).
inline u32 table_policy(int option, u64 m_cells) {
  if (option >= 0) return (u32)option;
  return m_cells * 8 <= (1ull << 30) ? 0u : 1u;
}

// The last error text of a context.  API calls write it on the caller's
// thread and a serving context's server thread writes it too, while
// fc5_last_error may read it from any thread: every access takes the lock.
class ErrText {
 public:
  ErrText& operator=(std::string s) {
    std::lock_guard<std::mutex> lk(mu_);
    s_ = std::move(s);
    return *this;
  }
  std::string get() const {
    std::lock_guard<std::mutex> lk(mu_);
    return s_;
  }

 private:
  mutable std::mutex mu_;
  std::string s_;
};

inline u64 inv_mod_2_64(u64 a) {  // a odd; Newton iteration doubles correct bits
  u64 x = a;
  for (int i = 0; i < 6; ++i) x *= 2 - a * x;
  return x;
}

}  // namespace fc5

using namespace fc5;

struct fc5_band {
  size_t n_in = 0, n_skipped = 0;
  std::vector<u8> cls_of;           // class of each caller index (0 = skipped)
  std::vector<int> slot_of;         // caller index -> index in units / units3
  // class 5
  std::vector<Cls5Unit> units;      // packed cls5 units
  std::vector<u32> uid;             // caller index of each packed cls5 unit
  std::vector<u32> irregular;       // cls5 units on the per-point 192-bit path
  DevUnit5* d_units = nullptr;
  u32 n_dev = 0, n_items = 0;
  PlainUnit* d_plain = nullptr;     // every packed cls5 unit, reference shape
  u32 n_plain = 0;
  u32 max_fmax2 = 0, max_lim = 0;
  // classes 2-4
  std::vector<Cls234Unit> units3;
  std::vector<u32> uid3;
  std::vector<u32> irregular3;
  DevUnit3* d_units3 = nullptr;
  u32 n_dev3 = 0, n_items3 = 0;
  Segment* d_segs = nullptr;        // launch segments over the band's two unit arrays
  u32 n_segs = 0;
  u32 max_lim3 = 0;
  // inputs and host images of the device units (re-numbered per rows-per-item)
  std::vector<fc5_unit> in_units;
  std::vector<u32> in_cls;
  std::vector<DevUnit5> dev_h;
  std::vector<DevUnit3> dev3_h;
  int rows = 0;  // rows per work item the item numbering was made for
  // captured batched pipeline
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<u64> graph_key;
};

struct fc5_ctx {
  int device = 0;
  int num_cu = 0;
  ErrText err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // options
  u64 hit_cap = 1u << 20;
  int use_gate = 1, verify = 1, stop_first = 0, blocks_per_cu = 0;
  int strict_gate = 0;  // 1: the strict residue gate before every xor query (fc5_strict.hpp)
  int sweep = 1;        // class-5 sweep of the batched kernel: 1 bit sieve, 0 per-point bitmap test
  int rows_per_item = 0, use_graph = 0;  // rows_per_item 0: by band size (auto_rows_per_item)
  Counters* h_ctr = nullptr;  // pinned read-back of the counters
  u64* h_nhits = nullptr;
  int load_policy = -1;  // cache policy of the xor cell gathers; -1: by table size (table_policy)
  // tables
  DeviceTable main, side;
  XorTable host_main;  // host copy when small enough (save, host verify)
  bool host_main_valid = false;
  u64 stream_load_min = (u64)4 << 30;  // option: fc5_table_load streams tables of more packed bytes
  // gate
  u32* d_gate = nullptr;
  u32* d_gate_ones = nullptr;
  GateWords gw{};
  // d table
  u32* d_ndg = nullptr;
  u32 ndg_K = 0, ndg_dmax = 0;
  // run buffers
  HitRec* d_hits = nullptr;
  u64 d_hits_cap = 0;
  Counters* d_ctr = nullptr;  // [cls] batched kernel, [8 + cls] plain kernels
  u16* d_p6 = nullptr;        // x^6 mod 504 (504 entries) then mod 247 (247 entries)
  fc5_result class_res[6]{};
  u64* d_nhits = nullptr;
  u32* d_item_ctr = nullptr;
  // last results
  std::vector<fc5_hit> hits;
  std::vector<fc5_solution> sols;
  // (6,1,7) class-5 twin
  int twin_tri = 1, twin_c4_div = 3, twin_plain = 0;
  int twin_tile = 0;  // 0: every c4 once; > 0: the reference launch's unbounded c4 tiles
  u32* d_tri = nullptr;               // triple-sum sets mod 504 (words 0..15) and 247 (16..23)
  TwinCounters* d_twin_ctr = nullptr;
  TwinCounters* h_twin_ctr = nullptr; // pinned read-back
  DevTwinCand* d_twin = nullptr;
  size_t d_twin_cap = 0;
  std::vector<fc5_twin_hit> twin_hits;
  std::vector<fc5_twin_solution> twin_sols;
  // exact sorted pair table (fc5_exact_api.cpp): n_exact entries value * 2^exact_shift + a
  u128* d_exact = nullptr;
  u64 n_exact = 0;
  int exact_N = 0;
  u32 exact_shift = 0;
  int verify_table = 0;  // option: answer the exact check from d_exact
  // resident serving loop (fc5_serve.cpp); while set, only fc5_serve_* may run
  struct fc5_server* server = nullptr;
  // partitioned probing (fc5_part_api.cpp): option probe_mode (0 a gather per
  // cell of every probe, 1 partitioned, -1 by table: partitioned for tables of
  // 2^28 cells or more with r >= 24), the emit buffer's capacity in hashes,
  // the buffers (the low-bits table belongs to the current main table), and
  // the last partitioned run's stage times and counts
  int probe_mode = -1;
  u64 part_records = (1ull << 32) - kEmitChunk;  // 1,048,575 emit chunks
  u64 part_hbm_limit = 0;  // bytes the buffers may take (0: no limit); beyond it, as on a failed allocation,
                           // probe_mode -1 probes with gathers and probe_mode 1 fails
  PartBufs* part = nullptr;
  double part_ms[3] = {0, 0, 0};  // emit launches, partition pipeline, identify launch (host clock)
  // event-timed pipeline stages summed over the pieces: first partition, fine
  // partition, lookup (cell 0), fine, lookup (cell 1), fine, last lookup, full queries
  double part_stage_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  u64 part_counts[4] = {0, 0, 0, 0};  // pieces, accepted hashes, full queries, emit chunks filled
};

// No C++ exception may cross the C ABI: every entry point catches what its
// body throws (a failed host allocation, above all) and turns it into an
// error code, with the text in the context when there is one.
inline int abi_exception(fc5_ctx* ctx) noexcept {
  int rc = FC5_E_HIP;
  const char* what = "internal error";
  try {
    throw;
  } catch (const std::bad_alloc&) {
    rc = FC5_E_NOMEM;
    what = "host allocation failed";
  } catch (...) {
  }
  if (ctx) {
    try {
      ctx->err = what;
    } catch (...) {
    }
  }
  return rc;
}

#define NOT_SERVING(ctx)                                                              \
  do {                                                                                \
    if ((ctx)->server) {                                                              \
      (ctx)->err = "the context is serving: only fc5_serve_* calls are allowed";      \
      return FC5_E_STATE;                                                             \
    }                                                                                 \
  } while (0)

#define FAIL(ctx, code, msg)  \
  do {                        \
    (ctx)->err = (msg);       \
    return (code);            \
  } while (0)

#define HIPCHK(ctx, expr)                                                              \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      (ctx)->err = std::string("HIP error ") + hipGetErrorString(_e) + " at " #expr;   \
      return FC5_E_HIP;                                                                \
    }                                                                                  \
  } while (0)

namespace fc5 {
// shared run-buffer helpers (fc5_api.cpp)
int ensure_ndg(fc5_ctx* ctx, u32 dmax);
int ensure_hits(fc5_ctx* ctx);
DevUnit5 dev_unit5(const Cls5Unit& U, u32 uid, u32 band);
DevUnit3 dev_unit3(const Cls234Unit& U, u32 uid, u32 band);
// work items: ceil(rows / E) per unit, numbered per arena.  A launch takes
// items from a 32-bit counter that every wave bumps once more before it
// stops, so a band holds at most kMaxItems; false when it holds more.
constexpr u64 kMaxItems = (1ull << 32) - (1ull << 20);
constexpr const char* kTooManyItems =
    "band too large for one launch (more than 2^32 - 2^20 work items at this rows_per_item): split it";
bool number_items(std::vector<DevUnit5>& dev, std::vector<DevUnit3>& dev3, int E, u32& n5, u32& n3);
// rows per work item for `rows` e- or d-rows in all: the largest of 32, 16, 8
// that still gives every wave of the persistent grid 8 items or more, else 4
int auto_rows_per_item(int num_cu, u64 rows, bool strict, bool sieve);
// serving loop (fc5_serve.cpp): stops and frees the server of a context
void serve_shutdown(fc5_ctx* ctx);
// Stage-1 units of classes in mask (bits 2..5) for B in [lo, hi] on the device,
// in the host emitter's order, on stream s (fc5_stage1_api.cpp)
int stage1_emit(fc5_ctx* ctx, hipStream_t s, u64 lo, u64 hi, unsigned mask, double lo_frac, double hi_frac,
                std::vector<u32>& cls, std::vector<fc5_unit>& units);
// exact check of filter survivors on the device: pairs[i] gets every
// a <= b <= targets[i].hi with a^6 + b^6 = targets[i].T, sorted -- by the pair
// search (k_pair_search), or from the exact sorted table when the option
// verify_table is set, which also marks in extra (if given) the targets equal
// to one of its extra values
// (on stream st; nullptr: the context's stream)
int device_pairs(fc5_ctx* ctx, const std::vector<PairTarget>& targets,
                 std::vector<std::vector<std::pair<u32, u32>>>& pairs, std::vector<u8>* extra = nullptr,
                 hipStream_t st = nullptr);
// pair jobs on the device through launch (the pair search or the exact-table
// lookup), into an output buffer that grows once when the first pass overflows
using PairLaunch = std::function<int(const PairJob*, u32, PairHit*, u32, u32*, hipStream_t)>;
int run_pair_jobs(fc5_ctx* ctx, const std::vector<PairJob>& jobs, const PairLaunch& launch, std::vector<PairHit>& out,
                  hipStream_t st = nullptr);
// the exact-table half of device_pairs (fc5_exact_api.cpp)
int exact_table_pairs(fc5_ctx* ctx, const std::vector<PairTarget>& targets,
                      std::vector<std::vector<std::pair<u32, u32>>>& pairs, std::vector<u8>* extra,
                      hipStream_t st = nullptr);
void free_exact(fc5_ctx* ctx);
// the allocation-free entry point's per-device set-up (fc5_noalloc.cpp): the
// device address of its scratch array, with `device` current; once per device
int noalloc_setup(int device);
// twin buffers (fc5_twin_api.cpp)
void free_twin(fc5_ctx* ctx);
// partitioned probing (fc5_part_api.cpp): whether a batched run of this
// context probes by partition, the run itself (counters and hits land where
// the persistent launch puts them; counters, hit slot and work counter zero
// on entry), and the buffers' release (also when the main table changes)
bool use_partitioned(const fc5_ctx* ctx);
int run_partitioned(fc5_ctx* ctx, const BatchedArgs& a, int grid, int rows);
void free_part(fc5_ctx* ctx);
}  // namespace fc5
