// Host engine: segment store, HBM segment cache, CPU/GPU split, merge and the
// analytic cost model (include/mordred_engine.h).
//
// Reference anchors: CacheManager.cu:34-96 (store + cache), 306-361 (cache /
// evict a segment); QueryOptimizer.cu:1674-1747 (segment groups by residency,
// operator device = GPU iff its columns are cached); QueryProcessing.cu:
// 1346-1469 (runQuery: dims, fact groups on their own streams / threads, merge);
// CPUProcessing.cu:439-1104 (the CPU operators); CostModel.cu:283-440 (cost);
// common.h:137-201 (column files), CacheManager.cu:154-184 and
// QueryOptimizer.cu:1640-1662, 1695-1706 (per-segment min/max, skipping).
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <array>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <climits>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mordred_costmodel.h"
#include "../../include/mordred_engine.h"
#include "../../include/mordred_hip.h"
#include "batch.hpp"
#include "stream_pack.hpp"

using mordred::kPackMaxBits;
using mordred::packed_words;

namespace mordred_host {

constexpr int kSegLog = 20;
constexpr int kSeg = 1 << kSegLog;
constexpr int kCpuChunk = 1 << 16;   // rows per CPU task
constexpr int kStreamBatch = 4;      // segments per streamed batch
constexpr int kMaxPrefixBatch = 16;  // segments per prefix-streamed batch, at most
static_assert(kMaxPrefixBatch <= mordred::kPackMaxSegs && kSegLog == mordred::kPackSegLog,
              "a prefix batch is unpacked in one launch of segments of the engine's size");
constexpr int kMixedChunk = 64;      // segments per GPU call of a mixed pipeline
constexpr int kSuffixTask = 8192;    // survivors per CPU task of a mixed pipeline
constexpr double kCacheLine = 64.0;  // CostModel.h:4

using Clock = std::chrono::steady_clock;
static double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// ------------------------------------------------------------ worker pool
// Persistent threads; run() hands out task indexes through an atomic counter
// and returns when every task has finished.
class Pool {
 public:
  explicit Pool(int n) {
    if (n <= 0) n = std::max(1u, std::thread::hardware_concurrency());
    for (int i = 0; i < n; i++) threads_.emplace_back([this, i] { loop(i); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  int size() const { return (int)threads_.size(); }
  void run(int ntasks, const std::function<void(int, int)>& fn) {
    if (ntasks <= 0) return;
    std::unique_lock<std::mutex> lk(mu_);
    job_ = &fn;
    ntasks_ = ntasks;
    next_.store(0);
    active_ = (int)threads_.size();
    gen_++;
    cv_.notify_all();
    done_.wait(lk, [this] { return active_ == 0; });
    job_ = nullptr;
  }

 private:
  void loop(int id) {
    long seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      const auto* fn = job_;
      const int nt = ntasks_;
      lk.unlock();
      for (int t = next_.fetch_add(1); t < nt; t = next_.fetch_add(1)) (*fn)(t, id);
      lk.lock();
      if (--active_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int, int)>* job_ = nullptr;
  int ntasks_ = 0, active_ = 0;
  long gen_ = 0;
  bool stop_ = false;
  std::atomic<int> next_{0};
};

// ------------------------------------------------------------ column store
struct Column {
  std::string name;
  long long n = 0;
  int nseg = 0;
  int* host = nullptr;          // pinned host copy of every value
  std::vector<int> slot;        // cache slot per segment, -1 = not cached
  int* d_idx = nullptr;         // device segment table (the reference col_idx)
  bool pinned_gpu = false;
  bool ranged = false;          // vmin / vmax computed
  int vmin = 0, vmax = 0;       // exact value range of the column
  // Per-segment [min, max] used for segment skipping: computed from the
  // values when the column is filled, or read from a minmax file that covers
  // them (CacheManager.cu:154-184).
  std::vector<int> smin, smax;
  std::vector<int> emin, emax;  // exact per-segment min / max of the values
  int clustered = -1;           // -1 unknown, else 1 if segments cover narrow ranges
  // the packed PCIe copy of prefix-streamed segments (stream_pack.hpp):
  // pinned, segment s at packed + s * packed_words(pbits), made on first use
  uint32_t* packed = nullptr;
  int pbits = 0;
  std::vector<char> packed_seg;
};

// Device buffer that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
};

}  // namespace mordred_host

using namespace mordred_host;

// Dictionary of a dimension's attribute tuples (batch dictionary form):
// per slot (key - min_key) % len the code of its row's tuple (0: no row),
// cb bits per code, on the device; tuple t (code t + 1) packed as in the
// attribute form (value - base + 1 per attribute, attrs in order).
struct DimDict {
  const Column* key = nullptr;
  std::vector<Column*> attrs;
  int len = 0, minkey = 0;
  bool ok = false;                // false: too many tuples or attributes wider than 32 bits
  int cb = 0;
  std::vector<unsigned> tuples;
  DevBuf codes;
  unsigned long long last_use = 0;
};

struct mordred_engine {
  int device = 0;
  int cache_segments = 0;
  int* cache = nullptr;
  std::vector<int> free_slots;
  std::unordered_map<std::string, std::unique_ptr<Column>> cols;
  std::unique_ptr<Pool> pool;
  hipStream_t s_main = nullptr;
  hipStream_t s_copy[2] = {nullptr, nullptr};
  short* d_identity = nullptr;  // device segment group 0, 1, 2, ...
  double bw_cpu = 42e9, bw_pci = 12e9;  // CostModel.h:5-6 (bytes per second)
  bool skipping = true;                 // segment skipping (main.cu:21)
  bool prefix_stream = true;            // COST_SPLIT prefix streaming
  double suffix_ns_per_line = 0;        // measured CPU-operator time per survivor line (0 = none yet)
  double build_ns_per_row = 0;          // measured CPU hash-table build time per dimension row
  int op_order = MORDRED_ORDER_GPU_FIRST;  // operator order of partly cached segment groups
  // MORDRED_ORDER_AUTO: per query and residency, the run time of each order
  // tried (0 = not yet)
  std::unordered_map<std::string, std::array<double, 3>> order_ms;
  hipStream_t s_order = nullptr;        // the calling thread's stream for ordered pipelines
  hipEvent_t ev_built = nullptr;        // recorded after a run's hash-table builds
  std::unordered_map<long long, double> seg_weight;  // (column id << 32 | segment) -> weight
  DevBuf d_res, d_sg, d_ht[4], d_mix, d_pmix, d_ord, d_res_ord;
  DevBuf d_pack[2];  // packed prefix segments per staging set (stream_pack.hpp)
  HostBuf h_stage;  // pinned scratch for segment groups / index updates
  // shared-scan batches: record tables, descriptor, result arena, the graph
  DevBuf d_rec[4], d_mk[4], d_desc, d_arena, d_status, d_touch;
  HostBuf h_desc, h_arena;
  int batch_forms = MORDRED_BATCH_FORM_ALL;     // dimension table forms batches may use
  bool batch_share = true;                      // identical queries of a batch share sums
  std::vector<std::unique_ptr<DimDict>> dicts;  // dictionary-form code tables
  unsigned long long dict_clock = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<char> gkey;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  mordred::BatchConfig batch_cfg{512, 1, 64 * 1024};
  bool batch_tuned = false;
  // request loop: persistent slot arenas and the learned cost model
  DevBuf d_serve_res, d_serve_touch;
  mordred_cost_model* cm = nullptr;
  // learned plan model for MORDRED_PLACE_COST_SPLIT (0 off, 1 use when
  // trusted, 2 also explore), trained on the runs' measured times
  mordred_cost_model* cm_place = nullptr;
  int place_mode = 0;
  int place_since_train = 0;
  long long place_trains = 0;
  long long learned_order_choices = 0;    // runs whose order the model changed
  long long learned_residency_calls = 0;  // place() calls weighted by the model
  unsigned long long place_rng = 0x2545F4914F6CDD1Dull;
  std::string err;
};

namespace {

constexpr int kMaxIdentity = 32767;

int fail(mordred_engine* e, const std::string& msg) {
  e->err = msg;
  return -1;
}

int hip_fail(mordred_engine* e, hipError_t rc, const char* what) {
  e->err = std::string(what) + ": " + hipGetErrorString(rc);
  return (int)rc;
}

#define HIP_OK(call)                                  \
  do {                                                \
    hipError_t rc_ = (call);                          \
    if (rc_ != hipSuccess) return hip_fail(e, rc_, #call); \
  } while (0)

Column* col(mordred_engine* e, const char* name) {
  if (!name) return nullptr;
  auto it = e->cols.find(name);
  return it == e->cols.end() ? nullptr : it->second.get();
}

int set_slot(mordred_engine* e, Column* c, int seg, int slot) {
  c->slot[seg] = slot;
  HIP_OK(hipMemcpy(c->d_idx + seg, &slot, sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

int cache_seg(mordred_engine* e, Column* c, int seg) {
  if (seg < 0 || seg >= c->nseg) return fail(e, "segment out of range");
  if (c->slot[seg] >= 0) return 0;
  if (e->free_slots.empty()) return fail(e, "segment cache full");
  const int slot = e->free_slots.back();
  e->free_slots.pop_back();
  const long long rows = std::min<long long>(kSeg, c->n - (long long)seg * kSeg);
  HIP_OK(hipMemcpy(e->cache + (size_t)slot * kSeg, c->host + (size_t)seg * kSeg, rows * sizeof(int),
                   hipMemcpyHostToDevice));
  return set_slot(e, c, seg, slot);
}

int evict_seg(mordred_engine* e, Column* c, int seg) {
  if (seg < 0 || seg >= c->nseg) return fail(e, "segment out of range");
  if (c->slot[seg] < 0 || c->pinned_gpu) return 0;
  e->free_slots.push_back(c->slot[seg]);
  return set_slot(e, c, seg, -1);
}

// ------------------------------------------------------------ column fill
void free_column(mordred_engine* e, Column* c) {
  if (e->device < 0) {
    std::free(c->host);
  } else {
    if (c->host) (void)hipHostFree(c->host);
    if (c->d_idx) (void)hipFree(c->d_idx);
    if (c->packed) (void)hipHostFree(c->packed);
  }
  c->host = nullptr;
  c->d_idx = nullptr;
  c->packed = nullptr;
  c->packed_seg.clear();
}

// Fill a column's host store segment by segment on the worker pool, from a
// caller buffer or (values == NULL) from a headerless little-endian int32
// file with pread (common.h:137-147 loadColumn; the host is little-endian, so
// the bytes are the values), and record each segment's min and max on the way:
// the statistics minmax.cpp:28-45 computes in an offline pass.
int fill_column(mordred_engine* e, Column* c, const int* values, int fd) {
  c->smin.assign(c->nseg, 0);
  c->smax.assign(c->nseg, 0);
  c->emin.assign(c->nseg, 0);
  c->emax.assign(c->nseg, 0);
  c->clustered = -1;
  std::atomic<int> short_read{0};
  e->pool->run(c->nseg, [&](int seg, int) {
    const long long b = (long long)seg * kSeg;
    const long long n = std::min<long long>(kSeg, c->n - b);
    int* dst = c->host + b;
    if (values) {
      std::memcpy(dst, values + b, (size_t)n * sizeof(int));
    } else {
      const size_t want = (size_t)n * sizeof(int);
      size_t done = 0;
      while (done < want) {
        const ssize_t got = pread(fd, reinterpret_cast<char*>(dst) + done, want - done,
                                  (off_t)((size_t)b * sizeof(int) + done));
        if (got <= 0) {
          short_read.store(1);
          return;
        }
        done += (size_t)got;
      }
    }
    int mn = dst[0], mx = dst[0];
    for (long long i = 1; i < n; i++) {
      mn = std::min(mn, dst[i]);
      mx = std::max(mx, dst[i]);
    }
    c->smin[seg] = c->emin[seg] = mn;
    c->smax[seg] = c->emax[seg] = mx;
  });
  if (short_read.load()) return fail(e, "column file shorter than n values: " + c->name);
  c->vmin = c->vmax = 0;
  for (int seg = 0; seg < c->nseg; seg++) {
    c->vmin = seg ? std::min(c->vmin, c->smin[seg]) : c->smin[seg];
    c->vmax = seg ? std::max(c->vmax, c->smax[seg]) : c->smax[seg];
  }
  c->ranged = true;
  return 0;
}

// New column of n values: host store, segment table, statistics, and with
// gpu_resident every segment cached.  values == NULL reads file descriptor fd.
int add_column(mordred_engine* e, const char* name, const int* values, int fd, long long n, int gpu_resident) {
  if (e->cols.count(name)) return fail(e, std::string("duplicate column ") + name);
  const int nseg = (int)((n + kSeg - 1) / kSeg);
  if (nseg > kMaxIdentity) return fail(e, "column longer than 32767 segments");
  auto owned = std::make_unique<Column>();
  Column* c = owned.get();
  c->name = name;
  c->n = n;
  c->nseg = nseg;
  c->slot.assign(nseg, -1);
  const size_t bytes = std::max<size_t>(1, (size_t)n * sizeof(int));
  if (e->device < 0) {
    c->host = static_cast<int*>(std::malloc(bytes));
    if (!c->host) return fail(e, "out of host memory");
  } else {
    HIP_OK(hipSetDevice(e->device));
    hipError_t rc = hipHostMalloc(reinterpret_cast<void**>(&c->host), bytes, hipHostMallocDefault);
    if (rc == hipSuccess) rc = hipMalloc(&c->d_idx, std::max(1, nseg) * sizeof(int));
    if (rc == hipSuccess)
      rc = hipMemcpy(c->d_idx, c->slot.data(), nseg * sizeof(int), hipMemcpyHostToDevice);
    if (rc != hipSuccess) {
      free_column(e, c);
      return hip_fail(e, rc, "column allocation");
    }
  }
  int rc = fill_column(e, c, values, fd);
  if (rc) {
    free_column(e, c);
    return rc;
  }
  e->cols[name] = std::move(owned);
  if (gpu_resident && e->device >= 0) {
    for (int seg = 0; seg < nseg; seg++)
      if ((rc = cache_seg(e, c, seg))) return rc;
    c->pinned_gpu = true;
  }
  return 0;
}

// ----------------------------------------------------------- resolved query
struct Resolved {
  const mordred_star_query* q = nullptr;
  std::vector<Column*> fact_cols;   // distinct fact columns the query reads
  Column* filt[2] = {nullptr, nullptr};
  Column* key[4] = {nullptr, nullptr, nullptr, nullptr};
  Column* dkey[4] = {nullptr, nullptr, nullptr, nullptr};
  Column* dfilt[4] = {nullptr, nullptr, nullptr, nullptr};
  Column* dval[4] = {nullptr, nullptr, nullptr, nullptr};
  const mordred_join* join[4] = {nullptr, nullptr, nullptr, nullptr};
  Column* a1 = nullptr;
  Column* a2 = nullptr;
  long long n = 0;
  int nseg = 0;
};

int resolve(mordred_engine* e, const mordred_star_query* q, Resolved& r) {
  if (!q || q->njoins < 0 || q->njoins > 4 || q->nfilters < 0 || q->nfilters > 2)
    return fail(e, "bad query shape");
  if (q->kind == MORDRED_KIND_GROUPBY && q->total_val <= 0) return fail(e, "total_val must be > 0");
  r.q = q;
  auto add_fact = [&](Column* c) {
    if (std::find(r.fact_cols.begin(), r.fact_cols.end(), c) == r.fact_cols.end()) r.fact_cols.push_back(c);
  };
  for (int i = 0; i < q->nfilters; i++) {
    if (!(r.filt[i] = col(e, q->filters[i].col))) return fail(e, "unknown fact filter column");
    add_fact(r.filt[i]);
  }
  for (int i = 0; i < q->njoins; i++) {
    const mordred_join& j = q->joins[i];
    if (j.slot < 1 || j.slot > 4 || r.join[j.slot - 1]) return fail(e, "join slots must be distinct in 1..4");
    // The aggregate operators probe only table 4 (filter_probe_aggr_GPU2,
    // GPUProcessing.h:1675-1776; filter_probe_aggr_CPU, CPUProcessing.cu:
    // 1628-1710): another join would be applied by some placements and not
    // by others, so such a query has no placement-independent answer.
    if (q->kind == MORDRED_KIND_AGGR && j.slot != 4)
      return fail(e, "aggregate queries join only table 4 (the date table)");
    const int k = j.slot - 1;
    r.join[k] = &j;
    if (!(r.key[k] = col(e, j.fact_key)) || !(r.dkey[k] = col(e, j.dim_key)))
      return fail(e, "unknown join column");
    if (j.filter_col && !(r.dfilt[k] = col(e, j.filter_col))) return fail(e, "unknown dimension filter column");
    if (j.group_col && !(r.dval[k] = col(e, j.group_col))) return fail(e, "unknown group column");
    if (j.dim_len <= 0) return fail(e, "dim_len must be > 0");
    add_fact(r.key[k]);
  }
  if (!(r.a1 = col(e, q->aggr1))) return fail(e, "unknown aggregate column");
  add_fact(r.a1);
  if (q->aggr2) {
    if (!(r.a2 = col(e, q->aggr2))) return fail(e, "unknown aggregate column");
    add_fact(r.a2);
  }
  r.n = r.fact_cols[0]->n;
  for (Column* c : r.fact_cols)
    if (c->n != r.n) return fail(e, "fact columns differ in length");
  r.nseg = r.fact_cols[0]->nseg;
  return 0;
}

bool segment_cached(const Resolved& r, int seg) {
  for (Column* c : r.fact_cols)
    if (c->slot[seg] < 0) return false;
  return true;
}

long long segment_rows(const Resolved& r, int seg) {
  return std::min<long long>(kSeg, r.n - (long long)seg * kSeg);
}

// ------------------------------------------------------------ segment skipping
// QueryOptimizer::checkPredicate (QueryOptimizer.cu:1640-1662) as used by
// groupBitmapSegmentTable (:1695-1706): a fact segment is skipped when, for one
// predicate of the query, the segment's [min, max] of the predicate's column
// misses the predicate's range.  The predicates are
//   - the fact filters (BETWEEN lo..hi);
//   - per probed join with a dimension filter, the foreign-key range that the
//     filter implies.  The reference writes that range by hand next to the
//     date filter (params->compare1/2[lo_orderdate], QueryOptimizer.cu:
//     2580-2595); here it is derived from the slots the dimension build fills.
// Why a skipped segment contributes nothing: a fact row with key x in
// [min_key, min_key + dim_len) probes slot x - min_key, which is filled only
// if a dimension row passing the filter has that slot; so if every key of the
// segment lies in [min_key, min_key + dim_len) and none in
// [min_key + first filled slot, min_key + last filled slot], no row of the
// segment survives the join.  A segment with keys outside the hashed range is
// never skipped on that join.
struct SkipPred {
  const Column* c;
  long long lo, hi;        // predicate range (empty if lo > hi)
  long long in_lo, in_hi;  // skip only segments whose range lies inside this
};

// A column whose segments each cover less than half of its value range on
// average: sorted or clustered, so that min/max can exclude segments.
bool clustered(Column* c) {
  if (c->clustered < 0) {
    const double span = (double)c->vmax - (double)c->vmin;
    double sum = 0;
    for (int seg = 0; seg < c->nseg; seg++) sum += (double)c->smax[seg] - (double)c->smin[seg];
    c->clustered = c->nseg > 1 && sum <= 0.5 * span * c->nseg;
  }
  return c->clustered == 1;
}

// First and last hash slot that join k's build fills (cpu_build's filter
// semantics, which equal the GPU build's); false if a key falls outside the
// hashed range.  Empty: first > last.
bool filled_slots(mordred_engine* e, const Resolved& r, int k, long long& first, long long& last) {
  const mordred_join& j = *r.join[k];
  const Column* key = r.dkey[k];
  const Column* fc = r.dfilt[k];
  const long long n = key->n;
  const int tasks = (int)((n + kCpuChunk - 1) / kCpuChunk);
  std::vector<long long> lo(tasks, LLONG_MAX), hi(tasks, LLONG_MIN);
  std::atomic<int> outside{0};
  e->pool->run(tasks, [&](int task, int) {
    const long long b = (long long)task * kCpuChunk, en = std::min(n, b + kCpuChunk);
    long long mn = LLONG_MAX, mx = LLONG_MIN;
    for (long long row = b; row < en; row++) {
      const int x = fc->host[row];
      bool pass;
      if (j.filter_mode == 1) pass = x >= j.c1 && x <= j.c2;
      else if (j.filter_mode == 2) pass = x == j.c1 || x == j.c2;
      else pass = x == j.c1;
      if (!pass) continue;
      const long long d = (long long)key->host[row] - j.min_key;
      if (d < 0) {
        outside.store(1);
        continue;
      }
      const long long slot = d % j.dim_len;
      mn = std::min(mn, slot);
      mx = std::max(mx, slot);
    }
    lo[task] = mn;
    hi[task] = mx;
  });
  if (outside.load()) return false;
  first = LLONG_MAX;
  last = LLONG_MIN;
  for (int t = 0; t < tasks; t++) {
    first = std::min(first, lo[t]);
    last = std::max(last, hi[t]);
  }
  if (first > last) first = 1, last = 0;
  return true;
}

// skip[s] = 1 for every fact segment the query's predicates exclude.
int skip_segments(mordred_engine* e, const Resolved& r, std::vector<char>& skip) {
  const mordred_star_query& q = *r.q;
  std::vector<SkipPred> preds;
  for (int i = 0; i < q.nfilters; i++)
    preds.push_back({r.filt[i], q.filters[i].lo, q.filters[i].hi, LLONG_MIN, LLONG_MAX});
  for (int k = 0; k < 4; k++) {
    // an aggregate query's GPU kernel probes the date table only
    // (filter_probe_aggr_GPU2), so only that join may exclude segments
    if (!r.join[k] || !r.dfilt[k] || (q.kind != MORDRED_KIND_GROUPBY && k != 3)) continue;
    if (!clustered(r.key[k])) continue;
    long long first, last;
    if (!filled_slots(e, r, k, first, last)) continue;
    const long long mk = r.join[k]->min_key;
    preds.push_back({r.key[k], mk + first, mk + last, mk, mk + r.join[k]->dim_len - 1});
  }
  skip.assign(r.nseg, 0);
  int n = 0;
  for (int seg = 0; seg < r.nseg; seg++) {
    for (const SkipPred& p : preds) {
      const long long mn = p.c->smin[seg], mx = p.c->smax[seg];
      if (mn < p.in_lo || mx > p.in_hi) continue;
      if (p.lo > p.hi || mx < p.lo || mn > p.hi) {
        skip[seg] = 1;
        n++;
        break;
      }
    }
  }
  return n;
}

// ------------------------------------------------------------ cost model
// CostModel::calculate_cost with every operator of the fact pipeline on the
// CPU (the GPU side costs 0): filters scan 4 B per row, probes scan 4 B and
// touch one cache line per row, the group-by reads n_aggr cache lines and one
// group line per row; each operator shrinks L by its selectivity
// (CostModel.cu:365-440).  Returns seconds.
struct Selectivity {
  double filt[2] = {1, 1};
  double join[4] = {1, 1, 1, 1};
};

double cpu_cost(const Resolved& r, const Selectivity& sel, double L, double bw_cpu) {
  double cost = 0;
  for (int i = 0; i < r.q->nfilters; i++) {
    cost += L * 4 / bw_cpu;
    L *= sel.filt[i];
  }
  for (int k = 0; k < 4; k++) {
    if (!r.join[k]) continue;
    cost += L * 4 / bw_cpu + L * kCacheLine / bw_cpu;
    L *= sel.join[k];
  }
  const int n_aggr = r.a2 ? 2 : 1;
  cost += L * kCacheLine * n_aggr / bw_cpu + L * kCacheLine / bw_cpu;
  return cost;
}

// CostModel::transfer_cost(M): every fact column of the segment over PCIe.
double transfer_cost(const Resolved& r, double L, double bw_pci) {
  return L * 4 * (double)r.fact_cols.size() / bw_pci;
}

// ------------------------------------------------------------ learned split
// Plans of MORDRED_PLACE_COST_SPLIT: stream k of the n non-resident segments
// over PCIe, the other n - k on the CPU operators.  Every plan gives the same
// exact result; only the time differs.  The analytic model times a plan as
// the slower of its two concurrent sides; the learned model (an fp32 MLP,
// mordred_costmodel.h) predicts the whole run's time from the features below
// and is trained on the measured times of the runs themselves.
// A segment sent to the GPU side costs t_gpu there and, when it is prefix
// streamed, t_gcpu of CPU operator time over its survivors; the CPU side
// also builds its hash tables first (t_fix).
double split_analytic_ms(int n, int k, double t_cpu, double t_gpu, double t_gcpu = 0, double t_fix = 0) {
  return std::max(t_fix + (n - k) * t_cpu + k * t_gcpu, k * t_gpu);
}

// The k of the two sides finishing together: k*t_gpu = t_fix + (n-k)*t_cpu + k*t_gcpu.
int split_balance(int n, double t_cpu, double t_gpu, double t_gcpu = 0, double t_fix = 0) {
  if (n <= 0 || t_gcpu >= t_cpu) return 0;
  return (int)std::min<double>(n, std::max(0.0, (t_fix + n * t_cpu) / (t_gpu + t_cpu - t_gcpu) + 0.5));
}

// order: MORDRED_ORDER_* of the mixed groups; mixed_frac: their share of the
// segments; mixed_ms: the analytic model's time for them.
void split_features(const mordred_engine* e, const Resolved& r, const Selectivity& sel, int cached, int n, int k,
                    double t_cpu, double t_str, double t_gcpu, double t_fix, int order, double mixed_frac,
                    double mixed_ms, float* f) {
  std::fill(f, f + MORDRED_COST_FEATURES, 0.f);
  const mordred_star_query& q = *r.q;
  int njoins = 0;
  for (int j = 0; j < 4; j++) njoins += r.join[j] != nullptr;
  f[0] = n > 0 ? (float)k / (float)n : 0.f;
  f[1] = (float)n / 1024.f;
  f[2] = (float)cached / 1024.f;
  for (int j = 0; j < 4; j++) f[3 + j] = r.join[j] ? (float)sel.join[j] : 1.f;
  f[7] = (float)sel.filt[0];
  f[8] = (float)sel.filt[1];
  f[9] = (float)r.fact_cols.size() / 8.f;
  f[10] = (float)njoins / 4.f;
  f[11] = (float)q.nfilters / 2.f;
  f[12] = (float)(std::log2((double)std::max(1, q.total_val)) / 16.0);
  f[13] = q.kind == MORDRED_KIND_GROUPBY ? 1.f : 0.f;
  f[14] = (float)std::log(std::max(t_cpu, 1e-3));
  f[15] = (float)std::log(std::max(t_str, 1e-3));
  f[16] = (float)std::log(std::max(split_analytic_ms(n, k, t_cpu, t_str, t_gcpu, t_fix), 1e-3));
  f[17] = (float)(std::log2((double)std::max(1, e->pool->size())) / 5.0);
  f[18] = (float)std::log(std::max(t_gcpu, 1e-3));
  if (mixed_frac > 0 && order >= 0 && order < 3) f[19 + order] = 1.f;
  f[22] = (float)mixed_frac;
  f[23] = (float)std::log(std::max(mixed_ms, 1e-3));
}

// The learned plan model may choose: trusted (more accurate than the analytic
// model on fresh runs), or, in MORDRED_LEARNED_FORCE, trained at least once.
bool place_model_ready(mordred_engine* e) {
  if (!e->cm_place || e->place_mode <= 0) return false;
  mordred_cost_model_stats cs{};
  if (mordred_cost_model_get_stats(e->cm_place, &cs) != 0) return false;
  return cs.trusted || (e->place_mode == MORDRED_LEARNED_FORCE && e->place_trains > 0);
}

// ------------------------------------------------------------ CPU operators
// The engine's CPU half (not the test oracle): same results as
// CPUProcessing.cu build_CPU (:1034-1104), filter_probe_group_by_CPU
// (:723-866) and filter_probe_aggr_CPU (:1628-1710), with per-thread dense
// accumulators instead of one atomic per row.
// Partly cached segments that share one residency pattern, run as a mixed
// pipeline (the reference's per-operator placement, QueryOptimizer.cu:
// 1718-1747, executeTableFact_v1 in QueryProcessing.cu:201): the fact filters
// and probes whose columns are cached run first on the GPU
// (filter_probe_GPU2 into offset lists), the survivors' row ids and dimension
// rows go to the host, and the CPU operators finish the rest of the pipeline.
struct MixedGroup {
  std::vector<int> segs;
  bool gpu_filter[2] = {false, false};
  bool gpu_probe[4] = {false, false, false, false};
  long long rows = 0;        // fact rows of the group
  long long survivors = 0;   // rows that left the GPU operators (all chunks)
};

// Cache lines the CPU operators of a mixed pipeline touch per survivor: one
// per remaining column or dimension value, two per remaining probe.
int mixed_lines(const Resolved& r, const MixedGroup& m) {
  int lines = 1 + (r.a2 ? 2 : 1);
  for (int i = 0; i < r.q->nfilters; i++) lines += !m.gpu_filter[i];
  for (int k = 0; k < 4; k++)
    if (r.join[k]) lines += m.gpu_probe[k] ? (r.dval[k] ? 1 : 0) : 2;
  return lines;
}

// Analytic time per segment of a mixed pipeline (CostModel.cu:283-363 costs
// the GPU operators at zero), as its two sides: `gpu` the transfers (the
// survivors' offsets and, with prefix set, the GPU operators' fact columns
// first), `cpu` the CPU operators, one random cache line per survivor and
// mixed_lines -- priced at the engine's measured time per line once a run
// has calibrated it, else at the CPU bandwidth.
struct MixedCost {
  double gpu = 0, cpu = 0;
  double total() const { return gpu + cpu; }
};
void column_range(Column* c);
// Bits a value of the column takes in the packed PCIe stream (values - vmin,
// stream_pack.hpp), or 0 when it crosses unpacked.
int stream_bits(Column* c) {
  column_range(c);
  const long long span = (long long)c->vmax - c->vmin;
  int b = 1;
  while (b < 32 && (1ll << b) <= span) b++;
  return b <= kPackMaxBits ? b : 0;
}
double stream_value_bytes(const mordred_engine* e, Column* c) {
  const int b = e->device >= 0 ? stream_bits(c) : 0;
  return b ? b / 8.0 : 4.0;
}

// The packed PCIe copy of segments `segs` of column c, made once per segment
// on the engine's workers.  The column crosses unpacked (c->pbits stays 0)
// when its values need more than kPackMaxBits bits or no pinned buffer can be
// had.
void pack_segments(mordred_engine* e, Column* c, const std::vector<int>& segs) {
  if (e->device < 0 || segs.empty()) return;
  const int bits = stream_bits(c);
  if (!bits) return;
  const long long pw = packed_words(bits);
  if (!c->packed) {
    void* p = nullptr;
    if (hipHostMalloc(&p, (size_t)c->nseg * (size_t)pw * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    c->packed = static_cast<uint32_t*>(p);
    c->pbits = bits;
    c->packed_seg.assign(c->nseg, 0);
  }
  std::vector<int> todo;
  for (int seg : segs)
    if (!c->packed_seg[seg]) todo.push_back(seg);
  e->pool->run((int)todo.size(), [&](int t, int) {
    const int seg = todo[t];
    uint32_t* w = c->packed + (size_t)seg * (size_t)pw;
    const long long n = std::min<long long>(kSeg, c->n - (long long)seg * kSeg);
    const int* v = c->host + (size_t)seg * kSeg;
    unsigned long long acc = 0;
    int have = 0;
    long long o = 0;
    for (long long j = 0; j < n; j++) {
      acc |= (unsigned long long)(uint32_t)(v[j] - c->vmin) << have;
      have += bits;
      while (have >= 32) {
        w[o++] = (uint32_t)acc;
        acc >>= 32;
        have -= 32;
      }
    }
    if (have > 0) w[o++] = (uint32_t)acc;
    while (o < pw) w[o++] = 0u;
  });
  for (int seg : todo) c->packed_seg[seg] = 1;
}

MixedCost mixed_cost(const mordred_engine* e, const Resolved& r, const Selectivity& sel, const MixedGroup& m,
                     bool prefix) {
  double surv = 1, col_bytes = 0;  // (the prefix columns' bytes per row over PCIe)
  int outs = 1;
  for (int i = 0; i < r.q->nfilters; i++)
    if (m.gpu_filter[i]) surv *= sel.filt[i], col_bytes += stream_value_bytes(e, r.filt[i]);
  for (int k = 0; k < 4; k++) {
    if (!r.join[k] || !m.gpu_probe[k]) continue;
    surv *= sel.join[k];
    col_bytes += stream_value_bytes(e, r.key[k]);
    outs += r.dval[k] != nullptr;
  }
  const double L = (double)kSeg * surv;
  const double per_line = e->suffix_ns_per_line > 0 ? e->suffix_ns_per_line * 1e-9 : kCacheLine / e->bw_cpu;
  MixedCost c;
  c.gpu = (L * 4.0 * outs + (prefix ? (double)kSeg * col_bytes : 0.0)) / e->bw_pci * 1e3;
  c.cpu = L * mixed_lines(r, m) * per_line * 1e3;
  return c;
}

// Survivors of one chunk of a mixed group, on the host.
struct MixedChunk {
  int group = 0;
  std::vector<int> lo;       // fact row ids
  std::vector<int> dim[4];   // dimension rows, for GPU-probed tables with a group value
};

// Chunks pass from the GPU side to the CPU side as each one lands on the
// host, so the CPU operators on one chunk overlap the GPU operators on the
// next.
struct MixedFeed {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<MixedChunk> ready;
  bool done = false, ok = true;
  void push(MixedChunk&& c) {
    {
      std::lock_guard<std::mutex> g(mu);
      ready.push_back(std::move(c));
    }
    cv.notify_one();
  }
  void finish(bool success) {
    {
      std::lock_guard<std::mutex> g(mu);
      if (done) return;
      done = true;
      ok = success;
    }
    cv.notify_all();
  }
  // false when no chunk is left (check ok)
  bool pop(MixedChunk& c) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return !ready.empty() || done; });
    if (ready.empty()) return false;
    c = std::move(ready.front());
    ready.pop_front();
    return true;
  }
};

// A host array allocated without initialisation (cpu_build clears it in
// parallel).
struct HostInts {
  std::unique_ptr<int[]> p;
  size_t n = 0;
  void alloc(size_t count) {
    p.reset(new int[count]);
    n = count;
  }
  int* data() const { return p.get(); }
  int& operator[](size_t i) const { return p[i]; }
};

struct CpuTables {
  HostInts ht[4];
  // one bit per slot, set when the slot holds a row: a probe that misses
  // reads only this (len/8 bytes, which stays in a core's cache) and not the
  // 8-byte slot of a table that may not
  std::vector<uint64_t> bits[4];
  // the CPU operators probe in this order: the probes commute, so the most
  // selective table goes first and the others see only the rows it keeps
  int order[4] = {0, 1, 2, 3};
  double frac[4] = {1, 1, 1, 1};  // fraction of slots that hold a row
  bool hit(int k, int h) const { return bits[k][(size_t)h >> 6] >> (h & 63) & 1u; }
};

// Probe order by estimated pass fraction, smallest first (ties: the smaller
// table first, whose random reads are more likely to hit the cache).
void set_probe_order(const Resolved& r, const Selectivity& sel, CpuTables& t) {
  std::stable_sort(t.order, t.order + 4, [&](int a, int b) {
    const double sa = r.join[a] ? sel.join[a] : 2.0, sb = r.join[b] ? sel.join[b] : 2.0;
    if (sa != sb) return sa < sb;
    const int la = r.join[a] ? r.join[a]->dim_len : 0, lb = r.join[b] ? r.join[b]->dim_len : 0;
    return la < lb;
  });
}

// The dimension filter of build_CPU (CPUProcessing.cu:1064-1069) with the
// GPU build's modes: 0 EQ, 1 BETWEEN, 2 EQ-or-EQ.
inline bool dim_pass(const mordred_join& j, int x) {
  if (j.filter_mode == 1) return x >= j.c1 && x <= j.c2;
  if (j.filter_mode == 2) return x == j.c1 || x == j.c2;
  if (j.filter_mode == 0) return x == j.c1;
  return true;
}

void cpu_build(mordred_engine* e, const Resolved& r, CpuTables& t) {
  // three passes, each one task list over all the query's tables so that
  // small tables do not leave workers idle: clear the slots, insert the
  // passing dimension rows, derive the slot bitmaps
  struct Task {
    int k;
    long long b, en;
  };
  constexpr long long kClear = 1 << 20;  // ints per clearing task
  std::vector<Task> clear, ins, bmp;
  for (int k = 0; k < 4; k++) {
    if (!r.join[k]) continue;
    const long long len = r.join[k]->dim_len;
    t.ht[k].alloc(2 * (size_t)len);
    t.bits[k].assign(((size_t)len + 63) / 64, 0);
    for (long long b = 0; b < 2 * len; b += kClear) clear.push_back({k, b, std::min(2 * len, b + kClear)});
    const long long n = r.dkey[k]->n;
    for (long long b = 0; b < n; b += kCpuChunk) ins.push_back({k, b, std::min(n, b + kCpuChunk)});
    const long long nw = (long long)t.bits[k].size();
    for (long long b = 0; b < nw; b += 4096) bmp.push_back({k, b, std::min(nw, b + 4096)});
  }
  e->pool->run((int)clear.size(), [&](int i, int) {
    std::memset(t.ht[clear[i].k].data() + clear[i].b, 0, (size_t)(clear[i].en - clear[i].b) * sizeof(int));
  });
  e->pool->run((int)ins.size(), [&](int i, int) {
    const int k = ins[i].k;
    const mordred_join& j = *r.join[k];
    int* ht = t.ht[k].data();
    const Column* key = r.dkey[k];
    const Column* fc = r.dfilt[k];
    const Column* vc = r.dval[k];
    for (long long row = ins[i].b; row < ins[i].en; row++) {
      if (fc && !dim_pass(j, fc->host[row])) continue;
      const int h = (key->host[row] - j.min_key) % j.dim_len;
      if (vc) ht[2 * (size_t)h] = vc->host[row];
      ht[2 * (size_t)h + 1] = (int)row + 1;
    }
  });
  e->pool->run((int)bmp.size(), [&](int i, int) {
    const int k = bmp[i].k;
    const size_t len = (size_t)r.join[k]->dim_len;
    const int* ht = t.ht[k].data();
    uint64_t* bw = t.bits[k].data();
    for (long long w = bmp[i].b; w < bmp[i].en; w++) {
      uint64_t b = 0;
      for (int x = 0; x < 64; x++) {
        const size_t h = (size_t)w * 64 + x;
        if (h < len && (ht[2 * h] | ht[2 * h + 1]) != 0) b |= 1ull << x;
      }
      bw[w] = b;
    }
  });
  for (int k = 0; k < 4; k++) {
    if (!r.join[k]) continue;
    long long set = 0;
    for (uint64_t w : t.bits[k]) set += __builtin_popcountll(w);
    t.frac[k] = r.join[k]->dim_len > 0 ? (double)set / (double)r.join[k]->dim_len : 1.0;
  }
}

struct CpuAcc {
  std::vector<long long> sum;
  std::vector<int> key;        // 4 per group, 0 = not seen with a non-zero key
  std::vector<uint8_t> touched;
  long long scalar = 0;
  std::vector<int> sel, selh;  // cpu_suffix scratch: kept positions, their slots (4 each)
};

void cpu_fact(mordred_engine* e, const Resolved& r, const CpuTables& t, const std::vector<int>& segs,
              int* res) {
  const mordred_star_query& q = *r.q;
  const int G = q.kind == MORDRED_KIND_GROUPBY ? q.total_val : 1;
  const int nw = e->pool->size();
  std::vector<CpuAcc> acc(nw);
  std::vector<std::pair<int, long long>> tasks;  // (segment, first row of chunk)
  for (int s : segs)
    for (long long b = (long long)s * kSeg; b < (long long)s * kSeg + segment_rows(r, s); b += kCpuChunk)
      tasks.emplace_back(s, b);
  const int* fcol[2] = {r.filt[0] ? r.filt[0]->host : nullptr, r.filt[1] ? r.filt[1]->host : nullptr};
  const int* kcol[4];
  const int* hts[4];
  int len[4], mk[4], mv[4], uq[4];
  for (int k = 0; k < 4; k++) {
    kcol[k] = r.join[k] ? r.key[k]->host : nullptr;
    hts[k] = r.join[k] ? t.ht[k].data() : nullptr;
    len[k] = r.join[k] ? r.join[k]->dim_len : 1;
    mk[k] = r.join[k] ? r.join[k]->min_key : 0;
    mv[k] = r.join[k] && r.join[k]->group_col ? r.join[k]->min_val : 0;
    uq[k] = r.join[k] && r.join[k]->group_col ? r.join[k]->unique_val : 0;
  }
  const int* a1 = r.a1->host;
  const int* a2 = r.a2 ? r.a2->host : nullptr;
  e->pool->run((int)tasks.size(), [&](int ti, int w) {
    CpuAcc& A = acc[w];
    if (q.kind == MORDRED_KIND_GROUPBY && A.sum.empty()) {
      A.sum.assign(G, 0);
      A.key.assign(4 * (size_t)G, 0);
      A.touched.assign(G, 0);
    }
    const long long b = tasks[ti].second;
    const long long en = (long long)tasks[ti].first * kSeg + segment_rows(r, tasks[ti].first);
    const long long stop = std::min(en, b + kCpuChunk);
    for (long long row = b; row < stop; row++) {
      bool ok = true;
      for (int i = 0; i < q.nfilters && ok; i++) {
        const int x = fcol[i][row];
        ok = x >= q.filters[i].lo && x <= q.filters[i].hi;
      }
      if (!ok) continue;
      int v[4] = {0, 0, 0, 0};
      for (int j = 0; j < 4 && ok; j++) {
        const int k = t.order[j];
        if (!kcol[k]) continue;
        const int h = (kcol[k][row] - mk[k]) % len[k];
        if (!t.hit(k, h)) ok = false;
        else v[k] = hts[k][2 * (size_t)h];
      }
      if (!ok) continue;
      const int x1 = a1[row], x2 = a2 ? a2[row] : 0;
      if (q.kind == MORDRED_KIND_AGGR) {
        A.scalar += (long long)(int)((unsigned)x1 * (unsigned)x2);
        continue;
      }
      const int h = ((v[0] - mv[0]) * uq[0] + (v[1] - mv[1]) * uq[1] + (v[2] - mv[2]) * uq[2] +
                     (v[3] - mv[3]) * uq[3]) % G;
      A.sum[h] += (long long)(int)((unsigned)x1 - (unsigned)x2);
      A.touched[h] = 1;
      for (int k = 0; k < 4; k++)
        if (v[k] != 0) A.key[4 * (size_t)h + k] = v[k];
    }
  });
  // combine: keys as the reference CPU path leaves them (non-zero keys only)
  if (q.kind == MORDRED_KIND_AGGR) {
    long long s = 0;
    for (auto& A : acc) s += A.scalar;
    reinterpret_cast<long long*>(res)[2] += s;
    return;
  }
  for (auto& A : acc) {
    if (A.sum.empty()) continue;
    for (int h = 0; h < G; h++) {
      if (!A.touched[h]) continue;
      for (int k = 0; k < 4; k++)
        if (A.key[4 * (size_t)h + k] != 0) res[6 * (size_t)h + k] = A.key[4 * (size_t)h + k];
      reinterpret_cast<long long*>(res + 6 * (size_t)h + 4)[0] += A.sum[h];
    }
  }
}

// The CPU part of a mixed pipeline (the offset-list CPU operators of
// CPUProcessing.cu, e.g. probe_group_by_CPU2 :301-438): for each survivor of
// the GPU operators, the remaining fact filters and probes, then the group-by
// or aggregate, into per-worker accumulators that persist across the chunks
// of a run (cpu_combine adds them into res once).  A table probed on the
// GPU contributes the group value of the dimension row the GPU found (the
// value its hash table holds for that row).
void cpu_suffix(mordred_engine* e, const Resolved& r, const CpuTables& t, const MixedGroup& m, const MixedChunk& c,
                std::vector<CpuAcc>& acc) {
  const mordred_star_query& q = *r.q;
  const int G = q.kind == MORDRED_KIND_GROUPBY ? q.total_val : 1;
  const long long total = (long long)c.lo.size();
  // at least two tasks per worker (a chunk of a few segments holds only tens
  // of thousands of survivors)
  const long long per = std::max<long long>(
      1024, std::min<long long>(kSuffixTask, (total + 2 * e->pool->size() - 1) / (2 * e->pool->size())));
  const int tasks = (int)((total + per - 1) / per);
  int len[4], mk[4], mv[4], uq[4];
  for (int k = 0; k < 4; k++) {
    len[k] = r.join[k] ? r.join[k]->dim_len : 1;
    mk[k] = r.join[k] ? r.join[k]->min_key : 0;
    mv[k] = r.join[k] && r.join[k]->group_col ? r.join[k]->min_val : 0;
    uq[k] = r.join[k] && r.join[k]->group_col ? r.join[k]->unique_val : 0;
  }
  // Two passes per task.  The first runs the CPU filters and tests the CPU
  // probes' slot bitmaps (small, cached), reading only the filter and key
  // columns at the survivors' rows (random rows: each a cache and TLB miss,
  // so they are prefetched kPrefetch rows ahead); it keeps the rows that pass
  // and their slots.  The second reads, for those rows only, the group values
  // (hash slots, dimension values) and the aggregate columns, again
  // prefetched ahead, and accumulates.
  constexpr int kPrefetch = 16;
  const int* scol[6];
  int nsc = 0;
  for (int f = 0; f < q.nfilters; f++)
    if (!m.gpu_filter[f]) scol[nsc++] = r.filt[f]->host;
  for (int k = 0; k < 4; k++)
    if (r.join[k] && !m.gpu_probe[k]) scol[nsc++] = r.key[k]->host;
  // (when the CPU probes keep most rows, the first pass saves nothing: one
  // pass then prefetches everything)
  double keep = 1;
  for (int k = 0; k < 4; k++)
    if (r.join[k] && !m.gpu_probe[k]) keep *= t.frac[k] * (double)len[k] / (double)std::max<long long>(1, r.dkey[k]->n);
  const bool two_pass = keep < 0.5;
  // single pass: the slots of CPU-probed tables above 1 MiB are prefetched
  // half the distance ahead, from keys prefetched earlier (a small table's
  // slot is in cache already)
  int big[4], nbig = 0;
  for (int k = 0; k < 4; k++)
    if (r.join[k] && !m.gpu_probe[k] && len[k] >= (1 << 17)) big[nbig++] = k;
  e->pool->run(tasks, [&](int task, int w) {
    CpuAcc& A = acc[w];
    if (q.kind == MORDRED_KIND_GROUPBY && A.sum.empty()) {
      A.sum.assign(G, 0);
      A.key.assign(4 * (size_t)G, 0);
      A.touched.assign(G, 0);
    }
    const long long b = (long long)task * per, en = std::min(total, b + per);
    A.sel.clear();
    A.selh.clear();
    if (!two_pass) {
      for (long long i = b; i < en; i++) {
        if (i + kPrefetch < en) {
          const long long prow = c.lo[i + kPrefetch];
          for (int j = 0; j < nsc; j++) __builtin_prefetch(scol[j] + prow);
          __builtin_prefetch(r.a1->host + prow);
          if (r.a2) __builtin_prefetch(r.a2->host + prow);
          for (int k = 0; k < 4; k++)
            if (m.gpu_probe[k] && r.dval[k]) __builtin_prefetch(r.dval[k]->host + c.dim[k][i + kPrefetch]);
        }
        if (nbig && i + kPrefetch / 2 < en) {
          const long long prow = c.lo[i + kPrefetch / 2];
          for (int j = 0; j < nbig; j++)
            __builtin_prefetch(t.ht[big[j]].data() +
                               2 * (size_t)((r.key[big[j]]->host[prow] - mk[big[j]]) % len[big[j]]));
        }
        const long long row = c.lo[i];
        bool ok = true;
        for (int f = 0; f < q.nfilters && ok; f++) {
          if (m.gpu_filter[f]) continue;
          const int x = r.filt[f]->host[row];
          ok = x >= q.filters[f].lo && x <= q.filters[f].hi;
        }
        int v[4] = {0, 0, 0, 0};
        for (int j = 0; j < 4 && ok; j++) {
          const int k = t.order[j];
          if (!r.join[k]) continue;
          if (m.gpu_probe[k]) {
            if (r.dval[k]) v[k] = r.dval[k]->host[c.dim[k][i]];
            continue;
          }
          const int hk = (r.key[k]->host[row] - mk[k]) % len[k];
          ok = t.hit(k, hk);
          if (ok) v[k] = t.ht[k][2 * (size_t)hk];
        }
        if (!ok) continue;
        const int x1 = r.a1->host[row], x2 = r.a2 ? r.a2->host[row] : 0;
        if (q.kind == MORDRED_KIND_AGGR) {
          A.scalar += (long long)(int)((unsigned)x1 * (unsigned)x2);
          continue;
        }
        const int h = ((v[0] - mv[0]) * uq[0] + (v[1] - mv[1]) * uq[1] + (v[2] - mv[2]) * uq[2] +
                       (v[3] - mv[3]) * uq[3]) % G;
        A.sum[h] += (long long)(int)((unsigned)x1 - (unsigned)x2);
        A.touched[h] = 1;
        for (int k = 0; k < 4; k++)
          if (v[k] != 0) A.key[4 * (size_t)h + k] = v[k];
      }
      return;
    }
    for (long long i = b; i < en; i++) {
      if (i + kPrefetch < en) {
        const long long prow = c.lo[i + kPrefetch];
        for (int j = 0; j < nsc; j++) __builtin_prefetch(scol[j] + prow);
      }
      const long long row = c.lo[i];
      bool ok = true;
      for (int f = 0; f < q.nfilters && ok; f++) {
        if (m.gpu_filter[f]) continue;
        const int x = r.filt[f]->host[row];
        ok = x >= q.filters[f].lo && x <= q.filters[f].hi;
      }
      int hs[4] = {0, 0, 0, 0};
      for (int j = 0; j < 4 && ok; j++) {
        const int k = t.order[j];
        if (!r.join[k] || m.gpu_probe[k]) continue;
        hs[k] = (r.key[k]->host[row] - mk[k]) % len[k];
        ok = t.hit(k, hs[k]);
      }
      if (!ok) continue;
      A.sel.push_back((int)(i - b));
      A.selh.insert(A.selh.end(), hs, hs + 4);
    }
    const int ns = (int)A.sel.size();
    for (int s2 = 0; s2 < ns; s2++) {
      if (s2 + kPrefetch < ns) {
        const long long pi = b + A.sel[s2 + kPrefetch];
        const long long prow = c.lo[pi];
        __builtin_prefetch(r.a1->host + prow);
        if (r.a2) __builtin_prefetch(r.a2->host + prow);
        for (int k = 0; k < 4; k++) {
          if (!r.join[k] || !r.dval[k]) continue;
          if (m.gpu_probe[k]) __builtin_prefetch(r.dval[k]->host + c.dim[k][pi]);
          else __builtin_prefetch(t.ht[k].data() + 2 * (size_t)A.selh[4 * (size_t)(s2 + kPrefetch) + k]);
        }
      }
      const long long i = b + A.sel[s2];
      const long long row = c.lo[i];
      int v[4] = {0, 0, 0, 0};
      for (int k = 0; k < 4; k++) {
        if (!r.join[k]) continue;
        if (m.gpu_probe[k]) {
          if (r.dval[k]) v[k] = r.dval[k]->host[c.dim[k][i]];
        } else {
          v[k] = t.ht[k][2 * (size_t)A.selh[4 * (size_t)s2 + k]];  // the value word of a hit slot
        }
      }
      const int x1 = r.a1->host[row], x2 = r.a2 ? r.a2->host[row] : 0;
      if (q.kind == MORDRED_KIND_AGGR) {
        A.scalar += (long long)(int)((unsigned)x1 * (unsigned)x2);
        continue;
      }
      const int h = ((v[0] - mv[0]) * uq[0] + (v[1] - mv[1]) * uq[1] + (v[2] - mv[2]) * uq[2] +
                     (v[3] - mv[3]) * uq[3]) % G;
      A.sum[h] += (long long)(int)((unsigned)x1 - (unsigned)x2);
      A.touched[h] = 1;
      for (int k = 0; k < 4; k++)
        if (v[k] != 0) A.key[4 * (size_t)h + k] = v[k];
    }
  });
}

// Adds per-worker accumulators into res (CPU convention: non-zero keys only).
void cpu_combine(const Resolved& r, std::vector<CpuAcc>& acc, int* res) {
  const mordred_star_query& q = *r.q;
  const int G = q.kind == MORDRED_KIND_GROUPBY ? q.total_val : 1;
  if (q.kind == MORDRED_KIND_AGGR) {
    long long sum = 0;
    for (auto& A : acc) sum += A.scalar;
    reinterpret_cast<long long*>(res)[2] += sum;
    return;
  }
  for (auto& A : acc) {
    if (A.sum.empty()) continue;
    for (int h = 0; h < G; h++) {
      if (!A.touched[h]) continue;
      for (int k = 0; k < 4; k++)
        if (A.key[4 * (size_t)h + k] != 0) res[6 * (size_t)h + k] = A.key[4 * (size_t)h + k];
      reinterpret_cast<long long*>(res + 6 * (size_t)h + 4)[0] += A.sum[h];
    }
  }
}

// merge (CPUProcessing.cu:1786-1819): a CPU key word stays unless it is 0,
// then the GPU word is taken; sums add.
void merge(int* cpu, const int* gpu, int groups) {
  for (int i = 0; i < groups; i++) {
    for (int k = 0; k < 4; k++)
      if (cpu[i * 6 + k] == 0) cpu[i * 6 + k] = gpu[i * 6 + k];
    reinterpret_cast<unsigned long long*>(cpu)[i * 3 + 2] += reinterpret_cast<const unsigned long long*>(gpu)[i * 3 + 2];
  }
}

// ------------------------------------------------------------ GPU side
struct GpuArgs {
  filterArgsGPU f{};
  probeArgsGPU p{};
  groupbyArgsGPU g{};
  bool has_filter = false;
};

// Kernel arguments of the fused fact kernel.  With `staged` set, every fact
// column's segment table is the staged one (fact column i at staged + i*nseg)
// instead of the column's own col_idx; dimension columns keep their own.
void gpu_args(const Resolved& r, int* const ht[4], int* staged, GpuArgs& a) {
  const mordred_star_query& q = *r.q;
  auto idx = [&](Column* c) -> int* {
    if (staged) {
      auto it = std::find(r.fact_cols.begin(), r.fact_cols.end(), c);
      if (it != r.fact_cols.end()) return staged + (size_t)(it - r.fact_cols.begin()) * r.nseg;
    }
    return c->d_idx;
  };
  if (q.nfilters > 0) {
    a.has_filter = true;
    a.f.filter_idx1 = idx(r.filt[0]);
    a.f.compare1 = q.filters[0].lo;
    a.f.compare2 = q.filters[0].hi;
    a.f.mode1 = 1;
    if (q.nfilters > 1) {
      a.f.filter_idx2 = idx(r.filt[1]);
      a.f.compare3 = q.filters[1].lo;
      a.f.compare4 = q.filters[1].hi;
      a.f.mode2 = 1;
    }
  }
  int** keys[4] = {&a.p.key_idx1, &a.p.key_idx2, &a.p.key_idx3, &a.p.key_idx4};
  int** hts[4] = {&a.p.ht1, &a.p.ht2, &a.p.ht3, &a.p.ht4};
  int* lens[4] = {&a.p.dim_len1, &a.p.dim_len2, &a.p.dim_len3, &a.p.dim_len4};
  int* mks[4] = {&a.p.min_key1, &a.p.min_key2, &a.p.min_key3, &a.p.min_key4};
  int** grp[4] = {&a.g.group_idx1, &a.g.group_idx2, &a.g.group_idx3, &a.g.group_idx4};
  int* mvs[4] = {&a.g.min_val1, &a.g.min_val2, &a.g.min_val3, &a.g.min_val4};
  int* uqs[4] = {&a.g.unique_val1, &a.g.unique_val2, &a.g.unique_val3, &a.g.unique_val4};
  for (int k = 0; k < 4; k++) {
    if (!r.join[k]) continue;
    *keys[k] = idx(r.key[k]);
    *hts[k] = ht[k];
    *lens[k] = r.join[k]->dim_len;
    *mks[k] = r.join[k]->min_key;
    if (r.dval[k]) {
      *grp[k] = r.dval[k]->d_idx;
      *mvs[k] = r.join[k]->min_val;
      *uqs[k] = r.join[k]->unique_val;
    }
  }
  a.g.aggr_idx1 = idx(r.a1);
  a.g.aggr_idx2 = r.a2 ? idx(r.a2) : nullptr;
  a.g.total_val = q.kind == MORDRED_KIND_GROUPBY ? q.total_val : 1;
}

int launch_fact(mordred_engine* e, const Resolved& r, const GpuArgs& a, int n, int* d_res, const short* d_sg,
                hipStream_t s) {
  int rc;
  if (r.q->kind == MORDRED_KIND_AGGR)
    rc = mordred_filter_probe_aggr_GPU2(e->cache, a.has_filter ? &a.f : nullptr, &a.p, &a.g, n, d_res, 0, d_sg,
                                        (mordred_stream_t)s);
  else
    rc = mordred_filter_probe_group_by_GPU2(e->cache, a.has_filter ? &a.f : nullptr, &a.p, &a.g, n, d_res, 0,
                                            d_sg, (mordred_stream_t)s);
  if (rc != 0) return hip_fail(e, (hipError_t)rc, "fused fact kernel");
  return 0;
}

// Rows of a segment group; the partial last segment of the table, if present,
// must be the group's last entry (QueryOptimizer.cu:1708-1714).
long long group_rows(const Resolved& r, const std::vector<int>& segs) {
  long long n = 0;
  for (int s : segs) n += segment_rows(r, s);
  return n;
}

void order_last_partial(const Resolved& r, std::vector<int>& segs) {
  const int last = r.nseg - 1;
  if (segment_rows(r, last) == kSeg) return;
  auto it = std::find(segs.begin(), segs.end(), last);
  if (it != segs.end()) {
    segs.erase(it);
    segs.push_back(last);
  }
}

// Pass fractions without the hash tables, so that a run can plan before its
// CPU builds: a dimension filter over up to 2^16 dimension rows spread evenly
// over the table (the fraction of rows the build keeps), a fact filter over
// the first 2^14 fact rows.
Selectivity estimate_selectivity(const Resolved& r) {
  Selectivity s;
  for (int k = 0; k < 4; k++) {
    if (!r.join[k] || !r.dfilt[k] || r.dfilt[k]->n <= 0) continue;
    const Column* fc = r.dfilt[k];
    const long long step = std::max<long long>(1, fc->n >> 16);
    long long seen = 0, pass = 0;
    for (long long row = 0; row < fc->n; row += step, seen++) pass += dim_pass(*r.join[k], fc->host[row]);
    s.join[k] = (double)pass / (double)seen;
  }
  // fact filters: a sample of the first segment
  const long long m = std::min<long long>(r.n, 1 << 14);
  for (int i = 0; i < r.q->nfilters; i++) {
    long long pass = 0;
    for (long long row = 0; row < m; row++) {
      const int x = r.filt[i]->host[row];
      pass += x >= r.q->filters[i].lo && x <= r.q->filters[i].hi;
    }
    s.filt[i] = m > 0 ? (double)pass / (double)m : 1.0;
  }
  return s;
}

// ------------------------------------------------------------ GPU half of a run
// Runs on its own host thread so that the CPU operators start at once: the
// HIP runtime may block an enqueue (hipMemcpyAsync of a streamed segment)
// until earlier commands drain, which on the calling thread would serialise
// the PCIe stream and the CPU group.  The builds and the cached segment group
// are issued immediately; the streamed segments once the split is known.
struct SplitPlan {
  std::vector<int> stream;          // segments streamed whole over PCIe
  std::vector<MixedGroup>* mixed;   // partly cached groups (owned by the caller of the run)
  std::vector<int> pstream;         // segments of which only pgroup's GPU operators' columns are streamed
  MixedGroup* pgroup;               // (owned by the caller of the run)
};

struct GpuPart {
  std::vector<int> gpu_segs;              // cached segment group
  std::future<SplitPlan> streamed;        // the split, from the CPU side
  MixedFeed* feed = nullptr;              // mixed groups' survivors, chunk by chunk
  std::vector<int> res;                   // partial result (res_ints)
  int rc = 0;
  double ms = 0;
  long long gpu_rows = 0, stream_rows = 0;
  int gpu_segments = 0, stream_segments = 0, prefix_segments = 0;
  long long prefix_rows = 0;
  long long packed_bytes = 0;  // prefix columns' bytes that crossed PCIe packed
  // the hash tables are built (e->ev_built recorded): 0, or -1 if the GPU
  // side failed first
  std::promise<int> built;
  bool built_set = false;
  void signal_built(int v) {
    if (built_set) return;
    built_set = true;
    built.set_value(v);
  }
};

// Ends the feed (as failed) on every exit that has not ended it.
struct FeedGuard {
  MixedFeed* f = nullptr;
  GpuPart* g = nullptr;
  ~FeedGuard() {
    if (f) f->finish(false);
    if (g) g->signal_built(-1);
  }
};

// Staging slots borrowed from the free list for one run; on every exit the
// engine's streams are drained before the slots go back.
struct SlotLoan {
  mordred_engine* e;
  std::vector<int> slots;
  ~SlotLoan() {
    if (slots.empty()) return;
    for (hipStream_t s : {e->s_main, e->s_copy[0], e->s_copy[1]}) (void)hipStreamSynchronize(s);
    for (int s : slots) e->free_slots.push_back(s);
  }
};

struct Events {
  hipEvent_t built = nullptr, tables = nullptr;
  ~Events() {
    if (built) (void)hipEventDestroy(built);
    if (tables) (void)hipEventDestroy(tables);
  }
};

inline size_t align64(size_t b) { return (b + 63) & ~size_t(63); }

// The GPU operators of a mixed pipeline over n rows of the segment group d_sg,
// on stream s: the reference's kernel for the operator set (filter_GPU2 for
// filters, probe_GPU2 for probes of any tables, filter_probe_GPU2 for filters
// plus the date probe, else filter_GPU2 then probe_GPU3 over its row ids).
// fidx / kidx are the segment tables of the filter and key columns (a
// column's own, or staged ones).  Survivors' row ids go to out.lo_off, the
// dimension rows of tables with a group value to out.dim_off*, their count to
// *d_total (zeroed here).
int mixed_gpu_ops(mordred_engine* e, const Resolved& r, int* const ht[4], const MixedGroup& m, int* const fidx[2],
                  int* const kidx[4], long long n, const short* d_sg, int* d_total, int* d_flo,
                  const offsetGPU& out, hipStream_t s) {
  filterArgsGPU f{};
  bool any_f = false;
  for (int i = 0; i < 2; i++) {
    if (!m.gpu_filter[i]) continue;
    any_f = true;
    if (i == 0) {
      f.filter_idx1 = fidx[0];
      f.compare1 = r.q->filters[0].lo;
      f.compare2 = r.q->filters[0].hi;
      f.mode1 = 1;
    } else {
      f.filter_idx2 = fidx[1];
      f.compare3 = r.q->filters[1].lo;
      f.compare4 = r.q->filters[1].hi;
      f.mode2 = 1;
    }
  }
  probeArgsGPU pa{};
  bool any_p = false, date_only = true;
  for (int k = 0; k < 4; k++) {
    if (!m.gpu_probe[k]) continue;
    any_p = true;
    date_only &= k == 3;
    const mordred_join& j = *r.join[k];
    switch (k) {
      case 0: pa.key_idx1 = kidx[0]; pa.ht1 = ht[0]; pa.dim_len1 = j.dim_len; pa.min_key1 = j.min_key; break;
      case 1: pa.key_idx2 = kidx[1]; pa.ht2 = ht[1]; pa.dim_len2 = j.dim_len; pa.min_key2 = j.min_key; break;
      case 2: pa.key_idx3 = kidx[2]; pa.ht3 = ht[2]; pa.dim_len3 = j.dim_len; pa.min_key3 = j.min_key; break;
      default: pa.key_idx4 = kidx[3]; pa.ht4 = ht[3]; pa.dim_len4 = j.dim_len; pa.min_key4 = j.min_key; break;
    }
  }
  HIP_OK(hipMemsetAsync(d_total, 0, sizeof(int), s));
  const mordred_stream_t ms = (mordred_stream_t)s;
  int krc;
  if (!any_p) {
    krc = mordred_filter_GPU2(e->cache, &f, out.lo_off, (int)n, d_total, 0, d_sg, ms);
  } else if (!any_f) {
    krc = mordred_probe_GPU2(e->cache, &pa, &out, (int)n, d_total, 0, d_sg, ms);
  } else if (date_only) {
    krc = mordred_filter_probe_GPU2(e->cache, &f, &pa, &out, (int)n, d_total, 0, d_sg, ms);
  } else {
    krc = mordred_filter_GPU2(e->cache, &f, d_flo, (int)n, d_total, 0, d_sg, ms);
    if (!krc) {
      int nf = 0;
      HIP_OK(hipMemcpyAsync(&nf, d_total, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      HIP_OK(hipMemsetAsync(d_total, 0, sizeof(int), s));
      offsetGPU in{};
      in.lo_off = d_flo;
      krc = mordred_probe_GPU3(e->cache, &in, &pa, &out, nf, d_total, ms);
    }
  }
  if (krc) return hip_fail(e, (hipError_t)krc, "mixed pipeline GPU operators");
  return 0;
}

// Copies a finished batch's survivors to the host (waits for stream s) and
// hands them to the feed.
int mixed_collect(mordred_engine* e, const Resolved& r, const MixedGroup& m, int group, const int* d_total,
                  const offsetGPU& out, hipStream_t s, MixedFeed* feed, long long& survivors) {
  int got = 0;
  HIP_OK(hipMemcpyAsync(&got, d_total, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  MixedChunk ch;
  ch.group = group;
  ch.lo.resize(got);
  if (got > 0) HIP_OK(hipMemcpyAsync(ch.lo.data(), out.lo_off, (size_t)got * sizeof(int), hipMemcpyDeviceToHost, s));
  int* const douts[4] = {out.dim_off1, out.dim_off2, out.dim_off3, out.dim_off4};
  for (int k = 0; k < 4; k++) {
    if (!m.gpu_probe[k] || !r.dval[k]) continue;
    ch.dim[k].resize(got);
    if (got > 0)
      HIP_OK(hipMemcpyAsync(ch.dim[k].data(), douts[k], (size_t)got * sizeof(int), hipMemcpyDeviceToHost, s));
  }
  HIP_OK(hipStreamSynchronize(s));
  survivors += got;
  feed->push(std::move(ch));
  return 0;
}

int gpu_part(mordred_engine* e, const Resolved& r, GpuPart& g, size_t res_ints) {
  FeedGuard feed_guard{g.feed, &g};  // every early return ends the feed as failed
  const auto t0 = Clock::now();
  HIP_OK(hipSetDevice(e->device));
  HIP_OK(e->d_res.reserve(res_ints * sizeof(int)));
  int* d_res = static_cast<int*>(e->d_res.p);
  HIP_OK(hipMemsetAsync(d_res, 0, res_ints * sizeof(int), e->s_main));
  int* ht[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int k = 0; k < 4; k++) {
    if (!r.join[k]) continue;
    const mordred_join& j = *r.join[k];
    // the engine's hash tables stay registered (mordred_probe_table_register),
    // so the fused calls of one run derive each compact probe table once;
    // each run's builds mark them stale
    const size_t ht_bytes = 2 * (size_t)j.dim_len * sizeof(int);
    if (e->d_ht[k].p && ht_bytes > e->d_ht[k].bytes)
      (void)mordred_probe_table_unregister(static_cast<const int*>(e->d_ht[k].p));
    HIP_OK(e->d_ht[k].reserve(ht_bytes));
    ht[k] = static_cast<int*>(e->d_ht[k].p);
    HIP_OK((hipError_t)mordred_probe_table_register(ht[k], j.dim_len));
    HIP_OK(hipMemsetAsync(ht[k], 0, ht_bytes, e->s_main));
    filterArgsGPU bf{};
    if (r.dfilt[k]) {
      bf.filter_idx1 = r.dfilt[k]->d_idx;
      bf.mode1 = j.filter_mode;
      bf.compare1 = j.c1;
      bf.compare2 = j.c2;
    }
    buildArgsGPU ba{r.dkey[k]->d_idx, r.dval[k] ? r.dval[k]->d_idx : nullptr, j.dim_len, j.min_key, 0};
    const int brc = mordred_build_GPU2(e->cache, r.dfilt[k] ? &bf : nullptr, &ba, (int)r.dkey[k]->n, ht[k], 0,
                                       e->d_identity, (mordred_stream_t)e->s_main);
    if (brc) return hip_fail(e, (hipError_t)brc, "mordred_build_GPU2");
  }
  Events ev;
  HIP_OK(hipEventCreateWithFlags(&ev.built, hipEventDisableTiming));
  HIP_OK(hipEventRecord(ev.built, e->s_main));
  HIP_OK(hipEventRecord(e->ev_built, e->s_main));
  g.signal_built(0);

  // pinned + device scratch: [cached group: nseg shorts][staged segment tables:
  // nc x nseg ints][one segment group per streamed batch]
  const size_t nseg = (size_t)r.nseg, nc = r.fact_cols.size();
  const size_t sg_bytes = align64(nseg * sizeof(short)), tab_bytes = align64(nc * nseg * sizeof(int));
  const size_t batch_bytes = align64(std::max(kStreamBatch, kMaxPrefixBatch) * sizeof(short));
  const size_t max_batches = nseg / kStreamBatch + 2;  // whole and prefix batches
  const size_t scratch = sg_bytes + tab_bytes + max_batches * batch_bytes;
  HIP_OK(e->h_stage.reserve(scratch));
  HIP_OK(e->d_sg.reserve(scratch));
  char* hs = static_cast<char*>(e->h_stage.p);
  char* ds = static_cast<char*>(e->d_sg.p);

  GpuArgs a;
  gpu_args(r, ht, nullptr, a);
  if (!g.gpu_segs.empty()) {
    order_last_partial(r, g.gpu_segs);
    short* h_sg = reinterpret_cast<short*>(hs);
    for (size_t i = 0; i < g.gpu_segs.size(); i++) h_sg[i] = (short)g.gpu_segs[i];
    HIP_OK(hipMemcpyAsync(ds, h_sg, g.gpu_segs.size() * sizeof(short), hipMemcpyHostToDevice, e->s_main));
    g.gpu_rows = group_rows(r, g.gpu_segs);
    const int frc = launch_fact(e, r, a, (int)g.gpu_rows, d_res, reinterpret_cast<short*>(ds), e->s_main);
    if (frc) return frc;
    g.gpu_segments = (int)g.gpu_segs.size();
  }

  SplitPlan plan = g.streamed.get();
  std::vector<int>& seg_order = plan.stream;
  std::vector<int>& pseg = plan.pstream;
  SlotLoan loan{e, {}};
  if (!seg_order.empty() || !pseg.empty()) {
    // Two staging sets of kStreamBatch segments x fact columns.  Batch i uses
    // set i % 2 on copy stream i % 2, so a set is refilled only after the same
    // stream has run the kernel of the batch before it.  The staged tables map
    // every streamed segment to the slot of its batch; other entries keep the
    // column's own slot (or 0) and are not read by the streamed launches.
    // Whole-segment batches come first, then the prefix batches, whose
    // segments carry only the columns of the prefix group's GPU operators.
    order_last_partial(r, seg_order);
    order_last_partial(r, pseg);
    struct Batch {
      bool prefix;
      size_t first, cnt;  // into `all`
    };
    std::vector<int> all(seg_order);
    all.insert(all.end(), pseg.begin(), pseg.end());
    // columns a prefix batch copies (those of the prefix group's GPU
    // operators); a set's kStreamBatch x nc slots then hold pbatch segments
    // of these npc columns
    std::vector<char> pcopy(nc, 0);
    std::vector<int> pidx(nc, -1);
    int npc = 0;
    if (!pseg.empty()) {
      const MixedGroup& m = *plan.pgroup;
      for (size_t c = 0; c < nc; c++) {
        for (int i = 0; i < 2; i++) pcopy[c] |= m.gpu_filter[i] && r.fact_cols[c] == r.filt[i];
        for (int k = 0; k < 4; k++) pcopy[c] |= m.gpu_probe[k] && r.fact_cols[c] == r.key[k];
        if (pcopy[c]) pidx[c] = npc++;
      }
    }
    const int pbatch = npc > 0 ? std::max(1, std::min<int>(kMaxPrefixBatch, kStreamBatch * (int)nc / npc)) : 1;
    std::vector<Batch> batches;
    for (size_t i = 0; i < seg_order.size(); i += kStreamBatch)
      batches.push_back({false, i, std::min<size_t>(kStreamBatch, seg_order.size() - i)});
    // prefix batches taper at the end (half of what is left, once that is
    // below two full batches): the last batch's kernels, survivor copy and
    // CPU operators, which nothing overlaps, then cover one segment
    for (size_t i = 0; i < pseg.size();) {
      const size_t left = pseg.size() - i;
      const size_t cnt = left >= 2 * (size_t)pbatch ? (size_t)pbatch : std::max<size_t>(1, (left + 1) / 2);
      batches.push_back({true, seg_order.size() + i, cnt});
      i += cnt;
    }
    for (size_t i = 0; i < 2 * kStreamBatch * nc; i++) {
      loan.slots.push_back(e->free_slots.back());
      e->free_slots.pop_back();
    }
    // staging slot of segment i of batch bi, column c: a column's segments
    // of one batch take consecutive loaned slots, so that a run of
    // consecutive segments is one copy when the loaned slots are consecutive
    auto slot_of = [&](size_t bi, size_t i, size_t c) -> int {
      const size_t set = bi & 1;
      if (batches[bi].prefix) return loan.slots[set * kStreamBatch * nc + (size_t)pidx[c] * pbatch + i];
      return loan.slots[set * kStreamBatch * nc + c * kStreamBatch + i];
    };
    int* h_tab = reinterpret_cast<int*>(hs + sg_bytes);
    int* d_tab = reinterpret_cast<int*>(ds + sg_bytes);
    for (size_t c = 0; c < nc; c++) {
      const std::vector<int>& mirror = r.fact_cols[c]->slot;
      for (size_t s = 0; s < nseg; s++) h_tab[c * nseg + s] = std::max(0, mirror[s]);
      for (size_t bi = 0; bi < batches.size(); bi++) {
        if (batches[bi].prefix && !pcopy[c]) continue;
        for (size_t i = 0; i < batches[bi].cnt; i++) h_tab[c * nseg + all[batches[bi].first + i]] = slot_of(bi, i, c);
      }
    }
    for (int b = 0; b < 2; b++) HIP_OK(hipStreamWaitEvent(e->s_copy[b], ev.built, 0));
    HIP_OK(hipMemcpyAsync(d_tab, h_tab, nc * nseg * sizeof(int), hipMemcpyHostToDevice, e->s_copy[0]));
    HIP_OK(hipEventCreateWithFlags(&ev.tables, hipEventDisableTiming));
    HIP_OK(hipEventRecord(ev.tables, e->s_copy[0]));
    HIP_OK(hipStreamWaitEvent(e->s_copy[1], ev.tables, 0));
    GpuArgs as;
    gpu_args(r, ht, d_tab, as);
    // prefix batches: the staged tables of the operators' columns, the columns
    // to copy, and per staging set an output area [row ids | dimension rows x 4
    // | filtered row ids] of pbatch segments, the count
    auto staged_of = [&](const Column* c) -> int* {
      for (size_t i = 0; i < nc; i++)
        if (r.fact_cols[i] == c) return d_tab + i * nseg;
      return nullptr;
    };
    int* pf[2] = {nullptr, nullptr};
    int* pk[4] = {nullptr, nullptr, nullptr, nullptr};
    const size_t pcap = (size_t)pbatch * kSeg;
    offsetGPU pout[2] = {};
    int* ptotal[2] = {nullptr, nullptr};
    int* pflo[2] = {nullptr, nullptr};
    if (!pseg.empty()) {
      const MixedGroup& m = *plan.pgroup;
      for (int i = 0; i < 2; i++)
        if (m.gpu_filter[i]) pf[i] = staged_of(r.filt[i]);
      for (int k = 0; k < 4; k++)
        if (m.gpu_probe[k]) pk[k] = staged_of(r.key[k]);
      const size_t per_set = 6 * pcap + 16;
      HIP_OK(e->d_pmix.reserve(2 * per_set * sizeof(int)));
      for (int set = 0; set < 2; set++) {
        int* base = static_cast<int*>(e->d_pmix.p) + set * per_set;
        pout[set].lo_off = base;
        int** douts[4] = {&pout[set].dim_off1, &pout[set].dim_off2, &pout[set].dim_off3, &pout[set].dim_off4};
        for (int k = 0; k < 4; k++)
          if (m.gpu_probe[k] && r.dval[k]) *douts[k] = base + (size_t)(k + 1) * pcap;
        pflo[set] = base + 5 * pcap;
        ptotal[set] = base + 6 * pcap;
      }
    }
    // device area of each staging set for the packed prefix columns: per
    // prefix column pbatch segments of the widest packed segment
    long long pack_stride = 0;
    for (size_t c = 0; c < nc; c++)
      if (pcopy[c] && r.fact_cols[c]->pbits > 0)
        pack_stride = std::max(pack_stride, (long long)pbatch * packed_words(r.fact_cols[c]->pbits));
    if (pack_stride > 0)
      for (int set = 0; set < 2; set++)
        HIP_OK(e->d_pack[set].reserve((size_t)npc * (size_t)pack_stride * sizeof(uint32_t)));
    char* hb = hs + sg_bytes + tab_bytes;
    char* db = ds + sg_bytes + tab_bytes;
    long long psurv = 0;
    int pending = -1;  // prefix batch whose survivors are still on the device
    for (size_t bi = 0; bi < batches.size(); bi++) {
      const int set = (int)(bi & 1);
      hipStream_t s = e->s_copy[set];
      const Batch& B = batches[bi];
      short* h_bsg = reinterpret_cast<short*>(hb + bi * batch_bytes);
      short* d_bsg = reinterpret_cast<short*>(db + bi * batch_bytes);
      long long rows = 0;
      for (size_t i = 0; i < B.cnt; i++) {
        const int seg = all[B.first + i];
        h_bsg[i] = (short)seg;
        rows += segment_rows(r, seg);
      }
      // one copy per run of segments that are consecutive both in the host
      // column and in the staging slots; a prefix batch's packed columns
      // cross packed and are unpacked into the staging slots
      for (size_t c = 0; c < nc; c++) {
        if (B.prefix && !pcopy[c]) continue;
        Column* const col = r.fact_cols[c];
        bool packed = B.prefix && col->pbits > 0;
        for (size_t i = 0; packed && i < B.cnt; i++) packed = col->packed_seg[all[B.first + i]] != 0;
        if (packed) {
          const long long pw = packed_words(col->pbits);
          uint32_t* dpk = static_cast<uint32_t*>(e->d_pack[set].p) + (size_t)pidx[c] * (size_t)pack_stride;
          size_t i0 = 0;
          while (i0 < B.cnt) {
            size_t i1 = i0 + 1;
            while (i1 < B.cnt && all[B.first + i1] == all[B.first + i1 - 1] + 1) i1++;
            HIP_OK(hipMemcpyAsync(dpk + i0 * (size_t)pw, col->packed + (size_t)all[B.first + i0] * (size_t)pw,
                                  (i1 - i0) * (size_t)pw * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            i0 = i1;
          }
          mordred::UnpackBatch u{};
          u.packed = dpk;
          u.cache = e->cache;
          u.bits = col->pbits;
          u.base = col->vmin;
          u.nseg = (int)B.cnt;
          for (size_t i = 0; i < B.cnt; i++) {
            u.slot[i] = slot_of(bi, i, c);
            u.rows[i] = (int)segment_rows(r, all[B.first + i]);
          }
          HIP_OK(mordred::unpack_segments(u, s));
          g.packed_bytes += (long long)B.cnt * pw * (long long)sizeof(uint32_t);
          continue;
        }
        size_t i0 = 0;
        while (i0 < B.cnt) {
          size_t i1 = i0 + 1;
          while (i1 < B.cnt && all[B.first + i1] == all[B.first + i1 - 1] + 1 &&
                 slot_of(bi, i1, c) == slot_of(bi, i1 - 1, c) + 1)
            i1++;
          const int seg0 = all[B.first + i0];
          long long n = 0;
          for (size_t i = i0; i < i1; i++) n += segment_rows(r, all[B.first + i]);
          HIP_OK(hipMemcpyAsync(e->cache + (size_t)slot_of(bi, i0, c) * kSeg, r.fact_cols[c]->host + (size_t)seg0 * kSeg,
                                (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
          i0 = i1;
        }
      }
      HIP_OK(hipMemcpyAsync(d_bsg, h_bsg, B.cnt * sizeof(short), hipMemcpyHostToDevice, s));
      if (!B.prefix) {
        const int frc = launch_fact(e, r, as, (int)rows, d_res, d_bsg, s);
        if (frc) return frc;
        g.stream_rows += rows;
        continue;
      }
      int rc = mixed_gpu_ops(e, r, ht, *plan.pgroup, pf, pk, rows, d_bsg, ptotal[set], pflo[set], pout[set], s);
      if (rc) return rc;
      g.prefix_rows += rows;
      // the previous prefix batch (other set) finishes while this one runs
      if (pending >= 0) {
        rc = mixed_collect(e, r, *plan.pgroup, -1, ptotal[pending], pout[pending], e->s_copy[pending], g.feed, psurv);
        if (rc) return rc;
      }
      pending = set;
    }
    if (pending >= 0) {
      const int rc =
          mixed_collect(e, r, *plan.pgroup, -1, ptotal[pending], pout[pending], e->s_copy[pending], g.feed, psurv);
      if (rc) return rc;
    }
    if (plan.pgroup) plan.pgroup->survivors = psurv;
    g.stream_segments = (int)seg_order.size();
    g.prefix_segments = (int)pseg.size();
  }
  // mixed groups: the GPU operators into offset lists, then to the host, in
  // chunks of at most kMixedChunk segments (bounded device scratch)
  if (plan.mixed && !plan.mixed->empty()) {
    // device scratch: [row ids | dimension rows of tables 0-3 | a row-id list
    // between a filter call and a probe call] x cap ints, the survivor count,
    // the chunk's segment list
    const size_t cap = (size_t)kMixedChunk * kSeg;
    HIP_OK(e->d_mix.reserve((6 * cap + 16) * sizeof(int) + kMixedChunk * sizeof(short)));
    int* dm = static_cast<int*>(e->d_mix.p);
    int* d_total = dm + 5 * cap;
    int* d_flo = dm + 5 * cap + 16;
    short* d_msg = reinterpret_cast<short*>(dm + 6 * cap + 16);
    for (size_t gi = 0; gi < plan.mixed->size(); gi++) {
      MixedGroup& m = (*plan.mixed)[gi];
      order_last_partial(r, m.segs);
      int* fidx[2] = {nullptr, nullptr};
      int* kidx[4] = {nullptr, nullptr, nullptr, nullptr};
      for (int i = 0; i < 2; i++)
        if (m.gpu_filter[i]) fidx[i] = r.filt[i]->d_idx;
      offsetGPU out{};
      out.lo_off = dm;
      int** douts[4] = {&out.dim_off1, &out.dim_off2, &out.dim_off3, &out.dim_off4};
      for (int k = 0; k < 4; k++) {
        if (!m.gpu_probe[k]) continue;
        kidx[k] = r.key[k]->d_idx;
        if (r.dval[k]) *douts[k] = dm + (size_t)(k + 1) * cap;
      }
      m.survivors = 0;
      for (size_t c0 = 0; c0 < m.segs.size(); c0 += kMixedChunk) {
        const size_t c1 = std::min(m.segs.size(), c0 + kMixedChunk);
        std::vector<short> sgl(m.segs.begin() + c0, m.segs.begin() + c1);
        long long rows = 0;
        for (size_t i = c0; i < c1; i++) rows += segment_rows(r, m.segs[i]);
        HIP_OK(hipMemcpyAsync(d_msg, sgl.data(), sgl.size() * sizeof(short), hipMemcpyHostToDevice, e->s_main));
        int rc = mixed_gpu_ops(e, r, ht, m, fidx, kidx, rows, d_msg, d_total, d_flo, out, e->s_main);
        if (!rc) rc = mixed_collect(e, r, m, (int)gi, d_total, out, e->s_main, g.feed, m.survivors);
        if (rc) return rc;
      }
    }
  }
  if (g.feed) g.feed->finish(true);
  for (hipStream_t s : {e->s_main, e->s_copy[0], e->s_copy[1]}) HIP_OK(hipStreamSynchronize(s));
  g.res.assign(res_ints, 0);
  HIP_OK(hipMemcpy(g.res.data(), d_res, res_ints * sizeof(int), hipMemcpyDeviceToHost));
  g.ms = ms_since(t0);
  return 0;
}

// ------------------------------------------------------------ operator orders
// A partly cached segment group run in the operator order of one of the
// reference's fact pipelines (QueryProcessing.cu:201-437 CPU-first, :676-914
// GPU-first with a switch back).  A stage runs a set of operators on one
// device over the rows the previous stage kept; between stages on different
// devices the row ids and the dimension rows found so far cross PCIe.  Runs
// on the calling thread (which owns the CPU pool) on its own stream, chunk by
// chunk, after the GPU side has built the hash tables.
constexpr int kOrderChunk = 16;  // segments per chunk (bounds the device lists)

struct OrderStage {
  bool gpu = false;
  bool filt[2] = {false, false};
  bool probe[4] = {false, false, false, false};
  bool group = false;
  bool any_filter() const { return filt[0] || filt[1]; }
  bool any_probe() const { return probe[0] || probe[1] || probe[2] || probe[3]; }
  bool empty() const { return !group && !any_filter() && !any_probe(); }
};

std::vector<OrderStage> order_stages(const Resolved& r, const MixedGroup& m, int order) {
  const int nf = r.q->nfilters;
  // group-by on the GPU iff its fact columns are cached in every segment
  bool ggpu = order != MORDRED_ORDER_GPU_FIRST;
  for (int s : m.segs) ggpu = ggpu && r.a1->slot[s] >= 0 && (!r.a2 || r.a2->slot[s] >= 0);
  OrderStage st[4];
  auto cpu_f = [&](OrderStage& x) { for (int f = 0; f < nf; f++) x.filt[f] = !m.gpu_filter[f]; };
  auto gpu_f = [&](OrderStage& x) { for (int f = 0; f < nf; f++) x.filt[f] = m.gpu_filter[f]; };
  auto cpu_p = [&](OrderStage& x) { for (int k = 0; k < 4; k++) x.probe[k] = r.join[k] && !m.gpu_probe[k]; };
  auto gpu_p = [&](OrderStage& x) { for (int k = 0; k < 4; k++) x.probe[k] = r.join[k] && m.gpu_probe[k]; };
  if (order == MORDRED_ORDER_CPU_FIRST) {
    cpu_f(st[0]);
    st[1].gpu = true, gpu_f(st[1]), gpu_p(st[1]);
    cpu_p(st[2]);
  } else if (order == MORDRED_ORDER_ALTERNATE) {
    st[0].gpu = true, gpu_f(st[0]);
    cpu_f(st[1]), cpu_p(st[1]);
    st[2].gpu = true, gpu_p(st[2]);
  } else {
    st[0].gpu = true, gpu_f(st[0]), gpu_p(st[0]);
    cpu_f(st[1]), cpu_p(st[1]);
  }
  st[3].gpu = ggpu;
  st[3].group = true;
  std::vector<OrderStage> out;
  for (const OrderStage& x : st) {
    if (x.empty()) continue;
    if (!out.empty() && out.back().gpu == x.gpu) {
      OrderStage& y = out.back();
      for (int f = 0; f < 2; f++) y.filt[f] |= x.filt[f];
      for (int k = 0; k < 4; k++) y.probe[k] |= x.probe[k];
      y.group |= x.group;
    } else {
      out.push_back(x);
    }
  }
  return out;
}

// The rows a stage hands to the next: every row of the chunk (implicit), or
// a list of fact row ids with the dimension rows of the probed tables that
// have a group value, on the host or on the device.
struct RowList {
  bool implicit = true, on_gpu = false;
  long long n = 0;
  bool dim[4] = {false, false, false, false};
  std::vector<int> lo, dims[4];  // host copy
  int set = 0;                   // device copy: list set 0 or 1
};

// Device lists: two sets of [row ids | dimension rows 0-3] x cap ints, a row-id
// list between a filter call and a probe call, the count, the segment list.
struct OrderBufs {
  size_t cap = 0;
  int* set[2] = {nullptr, nullptr};
  int* flo = nullptr;
  int* total = nullptr;
  short* sg = nullptr;
  offsetGPU off(int which, const bool dim[4]) const {
    offsetGPU o{};
    int* b = set[which];
    o.lo_off = b;
    int** d[4] = {&o.dim_off1, &o.dim_off2, &o.dim_off3, &o.dim_off4};
    for (int k = 0; k < 4; k++)
      if (dim[k]) *d[k] = b + (size_t)(k + 1) * cap;
    return o;
  }
};

int order_count(mordred_engine* e, const OrderBufs& B, hipStream_t s, long long& n) {
  int got = 0;
  HIP_OK(hipMemcpyAsync(&got, B.total, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  n = got;
  return 0;
}

// Moves a list to the other device (switch_device_fact).
int order_switch(mordred_engine* e, RowList& L, const OrderBufs& B, hipStream_t s) {
  const size_t bytes = (size_t)L.n * sizeof(int);
  if (L.on_gpu) {
    L.lo.resize(L.n);
    int* b = B.set[L.set];
    if (bytes) HIP_OK(hipMemcpyAsync(L.lo.data(), b, bytes, hipMemcpyDeviceToHost, s));
    for (int k = 0; k < 4; k++) {
      if (!L.dim[k]) continue;
      L.dims[k].resize(L.n);
      if (bytes)
        HIP_OK(hipMemcpyAsync(L.dims[k].data(), b + (size_t)(k + 1) * B.cap, bytes, hipMemcpyDeviceToHost, s));
    }
    HIP_OK(hipStreamSynchronize(s));
  } else {
    L.set = 0;
    int* b = B.set[0];
    if (bytes) HIP_OK(hipMemcpyAsync(b, L.lo.data(), bytes, hipMemcpyHostToDevice, s));
    for (int k = 0; k < 4; k++)
      if (L.dim[k] && bytes)
        HIP_OK(hipMemcpyAsync(b + (size_t)(k + 1) * B.cap, L.dims[k].data(), bytes, hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));  // the host lists are reused
  }
  L.on_gpu = !L.on_gpu;
  return 0;
}

// A GPU stage (the reference's offset-mode kernels).  A group-by stage adds
// into d_res; any other stage leaves a device list.
int order_gpu_stage(mordred_engine* e, const Resolved& r, int* const ht[4], const OrderStage& st, RowList& L,
                    long long rows, const OrderBufs& B, int* d_res, hipStream_t s) {
  const mordred_stream_t ms = (mordred_stream_t)s;
  GpuArgs a;
  gpu_args(r, ht, nullptr, a);
  filterArgsGPU f{};
  for (int i = 0; i < 2; i++) {
    if (!st.filt[i]) continue;
    if (i == 0) f.filter_idx1 = r.filt[0]->d_idx, f.compare1 = r.q->filters[0].lo, f.compare2 = r.q->filters[0].hi, f.mode1 = 1;
    else f.filter_idx2 = r.filt[1]->d_idx, f.compare3 = r.q->filters[1].lo, f.compare4 = r.q->filters[1].hi, f.mode2 = 1;
  }
  probeArgsGPU p{};
  int** pk[4] = {&p.key_idx1, &p.key_idx2, &p.key_idx3, &p.key_idx4};
  int** ph[4] = {&p.ht1, &p.ht2, &p.ht3, &p.ht4};
  int* pl[4] = {&p.dim_len1, &p.dim_len2, &p.dim_len3, &p.dim_len4};
  int* pm[4] = {&p.min_key1, &p.min_key2, &p.min_key3, &p.min_key4};
  for (int k = 0; k < 4; k++) {
    if (!st.probe[k]) continue;
    *pk[k] = r.key[k]->d_idx, *ph[k] = ht[k], *pl[k] = r.join[k]->dim_len, *pm[k] = r.join[k]->min_key;
  }
  bool od[4];
  for (int k = 0; k < 4; k++) od[k] = L.dim[k] || (st.probe[k] && r.dval[k]);
  int krc = 0;
  if (L.implicit) {
    if (st.group) return fail(e, "operator order: a group-by stage over whole segments");
    MixedGroup ops;
    std::copy(st.filt, st.filt + 2, ops.gpu_filter);
    std::copy(st.probe, st.probe + 4, ops.gpu_probe);
    int* fidx[2] = {f.filter_idx1, f.filter_idx2};
    int* kidx[4] = {p.key_idx1, p.key_idx2, p.key_idx3, p.key_idx4};
    const offsetGPU out = B.off(0, od);
    const int rc = mixed_gpu_ops(e, r, ht, ops, fidx, kidx, rows, B.sg, B.total, B.flo, out, s);
    if (rc) return rc;
    L.implicit = false, L.on_gpu = true, L.set = 0;
    std::copy(od, od + 4, L.dim);
    return order_count(e, B, s, L.n);
  }
  int cur = L.set;
  if (st.any_filter()) {
    for (int k = 0; k < 4; k++)
      if (L.dim[k]) return fail(e, "operator order: a GPU filter after a probe");
    HIP_OK(hipMemsetAsync(B.total, 0, sizeof(int), s));
    krc = mordred_filter_GPU3(e->cache, B.set[cur], &f, B.set[1 - cur], (int)L.n, B.total, ms);
    if (krc) return hip_fail(e, (hipError_t)krc, "operator order: filter_GPU3");
    cur = 1 - cur;
    if (int rc = order_count(e, B, s, L.n)) return rc;
  }
  const bool groupby = r.q->kind == MORDRED_KIND_GROUPBY;
  // probes: compaction, unless the group-by kernel probes them itself
  // (filter_probe_group_by_GPU3 probes every table; the aggregate kernels
  // only the date table)
  const bool fused = st.group && (groupby || (!st.probe[0] && !st.probe[1] && !st.probe[2]));
  if (st.any_probe() && !fused) {
    const offsetGPU in = B.off(cur, L.dim), out = B.off(1 - cur, od);
    HIP_OK(hipMemsetAsync(B.total, 0, sizeof(int), s));
    krc = mordred_probe_GPU3(e->cache, &in, &p, &out, (int)L.n, B.total, ms);
    if (krc) return hip_fail(e, (hipError_t)krc, "operator order: probe_GPU3");
    cur = 1 - cur;
    std::copy(od, od + 4, L.dim);
    if (int rc = order_count(e, B, s, L.n)) return rc;
  }
  L.set = cur;
  if (!st.group) return 0;
  const offsetGPU in = B.off(cur, L.dim);
  const bool probes_left = fused && st.any_probe();
  if (groupby) krc = mordred_filter_probe_group_by_GPU3(e->cache, &in, nullptr, probes_left ? &p : nullptr, &a.g,
                                                        (int)L.n, d_res, ms);
  else if (probes_left) krc = mordred_probe_aggr_GPU3(e->cache, &in, &p, &a.g, (int)L.n, d_res, ms);
  else krc = mordred_aggregationGPU(e->cache, in.lo_off, &a.g, (int)L.n, d_res, ms);
  if (krc) return hip_fail(e, (hipError_t)krc, "operator order: group-by kernel");
  L.n = 0;
  return 0;
}

// A CPU stage without the group-by: the stage's filters and probes over the
// list (or every row of segs), keeping the passing rows and dimension rows.
void order_cpu_compact(mordred_engine* e, const Resolved& r, const CpuTables& t, const OrderStage& st,
                       const std::vector<int>& segs, RowList& L) {
  const mordred_star_query& q = *r.q;
  bool od[4];
  for (int k = 0; k < 4; k++) od[k] = L.dim[k] || (st.probe[k] && r.dval[k]);
  // tasks: (first row, end row) of a segment chunk, or a range of the list
  std::vector<std::pair<long long, long long>> tasks;
  if (L.implicit) {
    for (int sg : segs)
      for (long long b = (long long)sg * kSeg; b < (long long)sg * kSeg + segment_rows(r, sg); b += kCpuChunk)
        tasks.push_back({b, std::min<long long>((long long)sg * kSeg + segment_rows(r, sg), b + kCpuChunk)});
  } else {
    const long long per = std::max<long long>(1024, (L.n + 2 * e->pool->size() - 1) / (2 * e->pool->size()));
    for (long long b = 0; b < L.n; b += per) tasks.push_back({b, std::min(L.n, b + per)});
  }
  struct Out {
    std::vector<int> lo, dims[4];
  };
  std::vector<Out> outs(tasks.size());
  e->pool->run((int)tasks.size(), [&](int task, int) {
    Out& o = outs[task];
    for (long long i = tasks[task].first; i < tasks[task].second; i++) {
      const long long row = L.implicit ? i : L.lo[i];
      bool ok = true;
      for (int fi = 0; fi < q.nfilters && ok; fi++) {
        if (!st.filt[fi]) continue;
        const int x = r.filt[fi]->host[row];
        ok = x >= q.filters[fi].lo && x <= q.filters[fi].hi;
      }
      int drow[4] = {0, 0, 0, 0};
      for (int o = 0; o < 4 && ok; o++) {
        const int k = t.order[o];
        if (!st.probe[k]) continue;
        const mordred_join& j = *r.join[k];
        const int h = (r.key[k]->host[row] - j.min_key) % j.dim_len;
        const int slot_row = t.ht[k][2 * (size_t)h + 1];  // dimension row + 1, 0 = empty
        ok = slot_row != 0;
        drow[k] = slot_row - 1;
      }
      if (!ok) continue;
      o.lo.push_back((int)row);
      for (int k = 0; k < 4; k++)
        if (od[k]) o.dims[k].push_back(st.probe[k] ? drow[k] : L.dims[k][i]);
    }
  });
  RowList R;
  R.implicit = false;
  std::copy(od, od + 4, R.dim);
  size_t n = 0;
  for (const Out& o : outs) n += o.lo.size();
  R.lo.reserve(n);
  for (int k = 0; k < 4; k++)
    if (od[k]) R.dims[k].reserve(n);
  for (Out& o : outs) {
    R.lo.insert(R.lo.end(), o.lo.begin(), o.lo.end());
    for (int k = 0; k < 4; k++)
      if (od[k]) R.dims[k].insert(R.dims[k].end(), o.dims[k].begin(), o.dims[k].end());
  }
  R.n = (long long)n;
  L = std::move(R);
}

// Runs the mixed groups in the engine's operator order.  CPU group-by stages
// accumulate into acc, GPU ones into d_res (zeroed here).
int order_groups(mordred_engine* e, const Resolved& r, const CpuTables& t, std::vector<MixedGroup>& mixed, int order,
                 int* d_res, size_t res_ints, std::vector<CpuAcc>& acc, mordred_run_stats& stats) {
  HIP_OK(hipSetDevice(e->device));
  hipStream_t s = e->s_order;
  HIP_OK(hipStreamWaitEvent(s, e->ev_built, 0));
  HIP_OK(hipMemsetAsync(d_res, 0, res_ints * sizeof(int), s));
  int* ht[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int k = 0; k < 4; k++)
    if (r.join[k]) ht[k] = static_cast<int*>(e->d_ht[k].p);
  OrderBufs B;
  B.cap = (size_t)kOrderChunk * kSeg;
  HIP_OK(e->d_ord.reserve((11 * B.cap + 16) * sizeof(int) + kOrderChunk * sizeof(short)));
  int* base = static_cast<int*>(e->d_ord.p);
  B.set[0] = base;
  B.set[1] = base + 5 * B.cap;
  B.flo = base + 10 * B.cap;
  B.total = base + 11 * B.cap;
  B.sg = reinterpret_cast<short*>(base + 11 * B.cap + 16);
  for (MixedGroup& m : mixed) {
    order_last_partial(r, m.segs);
    const std::vector<OrderStage> stages = order_stages(r, m, order);
    if (stages.empty() || !stages.back().group) return fail(e, "operator order: no group-by stage");
    m.survivors = 0;
    for (size_t c0 = 0; c0 < m.segs.size(); c0 += kOrderChunk) {
      const size_t c1 = std::min(m.segs.size(), c0 + kOrderChunk);
      const std::vector<int> segs(m.segs.begin() + c0, m.segs.begin() + c1);
      const std::vector<short> sgl(segs.begin(), segs.end());
      const long long rows = group_rows(r, segs);
      HIP_OK(hipMemcpyAsync(B.sg, sgl.data(), sgl.size() * sizeof(short), hipMemcpyHostToDevice, s));
      HIP_OK(hipStreamSynchronize(s));  // sgl is pageable and goes out of scope with the chunk
      RowList L;
      L.n = rows;
      for (const OrderStage& st : stages) {
        if (!L.implicit && L.on_gpu != st.gpu) {
          if (int rc = order_switch(e, L, B, s)) return rc;
          stats.order_switches++;
        }
        if (st.group) m.survivors += L.n;
        if (st.gpu) {
          if (int rc = order_gpu_stage(e, r, ht, st, L, rows, B, d_res, s)) return rc;
        } else if (st.group) {
          if (L.implicit) return fail(e, "operator order: a CPU group-by stage over whole segments");
          // the CPU operators of cpu_suffix: everything before this stage is done
          MixedGroup done;
          for (int f = 0; f < 2; f++) done.gpu_filter[f] = !st.filt[f];
          for (int k = 0; k < 4; k++) done.gpu_probe[k] = r.join[k] && !st.probe[k];
          MixedChunk ch;
          ch.lo = std::move(L.lo);
          for (int k = 0; k < 4; k++)
            if (L.dim[k]) ch.dim[k] = std::move(L.dims[k]);
          cpu_suffix(e, r, t, done, ch, acc);
        } else {
          order_cpu_compact(e, r, t, st, segs, L);
        }
      }
    }
  }
  HIP_OK(hipStreamSynchronize(s));
  return 0;
}

// Key of a query and the residency patterns of its partly cached groups
// (MORDRED_ORDER_AUTO keeps one choice per key).
std::string order_signature(const Resolved& r, const std::vector<MixedGroup>& mixed) {
  const mordred_star_query& q = *r.q;
  std::string k;
  auto add = [&](const char* s) {
    k += s ? s : "-";
    k += '|';
  };
  auto addi = [&](long long v) {
    k += std::to_string(v);
    k += '|';
  };
  addi(q.kind), addi(q.total_val), add(q.aggr1), add(q.aggr2);
  for (int i = 0; i < q.nfilters; i++) add(q.filters[i].col), addi(q.filters[i].lo), addi(q.filters[i].hi);
  for (int i = 0; i < q.njoins; i++) {
    const mordred_join& j = q.joins[i];
    addi(j.slot), add(j.fact_key), add(j.dim_key), add(j.filter_col), addi(j.filter_mode), addi(j.c1), addi(j.c2),
        add(j.group_col), addi(j.min_val), addi(j.unique_val);
  }
  std::vector<int> pats;
  for (const MixedGroup& m : mixed) {
    int p = 0;
    for (int f = 0; f < 2; f++) p |= m.gpu_filter[f] << f;
    for (int j = 0; j < 4; j++) p |= m.gpu_probe[j] << (2 + j);
    pats.push_back(p);
  }
  std::sort(pats.begin(), pats.end());
  for (int p : pats) addi(p);
  return k;
}

// ------------------------------------------------------------ shared-scan batches
// Host side of batch.hpp: resolve the queries, lay out the descriptor, the
// LDS and the result arena, capture / replay the graph.
struct BatchPlan {
  mordred::BatchDesc d;
  long long n = 0;
  std::vector<long long> res_off;  // ints
  std::vector<int> G;
  int lds_queries = 0, global_queries = 0, dead = 0, lds32_queries = 0, keys_in_loop = 0;
  int shared = 0;  // identical queries that take another's sums
  int cols_read = 0;  // distinct fact columns the scan reads (algorithmic bytes: 4 per row each)
};

void column_range(Column* c) {
  if (c->ranged) return;
  int mn = 0, mx = 0;
  if (c->n > 0) {
    mn = mx = c->host[0];
    for (long long i = 1; i < c->n; i++) {
      mn = std::min(mn, c->host[i]);
      mx = std::max(mx, c->host[i]);
    }
  }
  c->vmin = mn;
  c->vmax = mx;
  c->ranged = true;
}

// Bit layout of packed attributes (attribute and dictionary forms): per
// attribute value - vmin + 1 in the fewest bits holding its span + 1, from
// bit 0 in order.  Returns the total bits.
int attr_layout(const std::vector<Column*>& attrs, int* off, int* bits) {
  int o = 0;
  for (size_t a = 0; a < attrs.size(); a++) {
    column_range(attrs[a]);
    const long long span = (long long)attrs[a]->vmax - attrs[a]->vmin + 2;
    int b = 1;
    while ((1ll << b) < span) b++;
    off[a] = o;
    bits[a] = b;
    o += b;
  }
  return o;
}

// The dictionary of a dimension's attribute tuples over attrs (or a cached
// one over a superset of them), built on the host from the columns' host
// copies and kept (at most 16, the least recently used dropped).  A new
// dictionary covers the attributes of the cached one for the same key too
// (while at most 3), so a batch stream converges on one per dimension.
// *out = nullptr when the rows have more than kBatchDictMax distinct tuples.
int dim_dict(mordred_engine* e, Column* key, const std::vector<Column*>& attrs, int len, int minkey,
             const DimDict** out) {
  using namespace mordred;
  *out = nullptr;
  DimDict* wider = nullptr;
  for (auto& d : e->dicts) {
    if (d->key != key || d->len != len || d->minkey != minkey) continue;
    bool covers = true;
    for (Column* c : attrs) covers &= std::find(d->attrs.begin(), d->attrs.end(), c) != d->attrs.end();
    if (covers) {
      d->last_use = ++e->dict_clock;
      if (d->ok) *out = d.get();
      return 0;
    }
    wider = d.get();
  }
  auto d = std::make_unique<DimDict>();
  d->key = key;
  d->len = len;
  d->minkey = minkey;
  d->attrs = attrs;
  if (wider && wider->ok)
    for (Column* c : wider->attrs)
      if (std::find(d->attrs.begin(), d->attrs.end(), c) == d->attrs.end() && d->attrs.size() < (size_t)kBatchMaxAttr)
        d->attrs.push_back(c);
  d->last_use = ++e->dict_clock;
  int off[kBatchMaxAttr], bits[kBatchMaxAttr];
  const int total = attr_layout(d->attrs, off, bits);
  const int na = (int)d->attrs.size();
  std::vector<unsigned> slot_code((size_t)len + 1, 0);
  std::unordered_map<unsigned, int> code;
  d->ok = total <= 32;
  for (long long i = 0; d->ok && i < key->n; i++) {
    const int v = key->host[i] - minkey;
    const int sl = v % len;  // the probe's slot (C remainder, as phash)
    if (sl < 0) continue;
    unsigned w = 0;
    for (int a = 0; a < na; a++)
      w |= (((unsigned)(d->attrs[a]->host[i] - d->attrs[a]->vmin) + 1u) & ((bits[a] == 32 ? 0u : (1u << bits[a])) - 1u))
           << off[a];
    auto it = code.find(w);
    if (it == code.end()) {
      if ((int)d->tuples.size() == kBatchDictMax) {
        d->ok = false;
        break;
      }
      it = code.emplace(w, (int)d->tuples.size() + 1).first;
      d->tuples.push_back(w);
    }
    slot_code[sl] = (unsigned)it->second;
  }
  if (d->ok) {
    int cb = 1;
    while ((1ll << cb) < (long long)d->tuples.size() + 1) cb++;
    d->cb = cb;
    // codes at bit slot * cb; 8 bytes of slack for the scan's 8-byte loads
    const size_t bytes = ((((size_t)len + 1) * cb + 7) / 8 + 8 + 15) & ~(size_t)15;
    std::vector<unsigned char> buf(bytes, 0);
    for (size_t sl = 0; sl <= (size_t)len; sl++) {
      const unsigned c = slot_code[sl];
      if (!c) continue;
      const size_t bit = sl * cb;
      for (int b = 0; b < cb; b++)
        if (c >> b & 1u) buf[(bit + b) >> 3] |= (unsigned char)(1u << ((bit + b) & 7));
    }
    HIP_OK(d->codes.reserve(bytes));
    HIP_OK(hipMemcpy(d->codes.p, buf.data(), bytes, hipMemcpyHostToDevice));
  }
  if (e->dicts.size() >= 16) {
    auto lru = std::min_element(e->dicts.begin(), e->dicts.end(),
                                [](const auto& x, const auto& y) { return x->last_use < y->last_use; });
    // (a launch in flight may read its codes)
    HIP_OK(hipStreamSynchronize(e->s_main));
    e->dicts.erase(lru);
  }
  e->dicts.push_back(std::move(d));
  if (e->dicts.back()->ok) *out = e->dicts.back().get();
  return 0;
}

// Whether query i probes table k: every join of a group-by query; only the
// date table (slot 4) for an aggregate query, like filter_probe_aggr_GPU2
// (GPUProcessing.h:1675-1776), which probes ht4 alone.
bool batch_probes(const mordred_star_query& q, const Resolved& r, int k) {
  return r.join[k] && (q.kind == MORDRED_KIND_GROUPBY || k == 3);
}

// Slots of a batch: query i holds bit i of every mask.  An empty slot
// (qp[i] == NULL) takes no part.  With `fixed` set, slot i's result rows sit
// at a fixed stride of the given arena (a serving loop keeps them across
// launches); otherwise they are packed into the engine's batch arena.
struct SlotArenas {
  int* res = nullptr;
  long long res_stride = 0;    // ints per slot
  unsigned* touch = nullptr;   // touched bitmaps
  int touch_stride = 0;        // words per slot
};

struct QueryEst {
  double sel = 1.0;
  int G = 1;
  int njoins = 0;
  int dims = 0;
};

double dim_pass_fraction(const Resolved& r, int k) {
  const Column* f = r.dfilt[k];
  if (!f) return 1.0;
  const mordred_join& j = *r.join[k];
  const long long n = r.dkey[k]->n;
  const long long step = std::max<long long>(1, n / 4096);
  long long seen = 0, pass = 0;
  for (long long row = 0; row < n; row += step, seen++) {
    const int x = f->host[row];
    const bool ok = j.filter_mode == 0 ? x == j.c1 : j.filter_mode == 1 ? (x >= j.c1 && x <= j.c2)
                                                                       : (x == j.c1 || x == j.c2);
    pass += ok;
  }
  return seen ? (double)pass / (double)seen : 1.0;
}

int estimate_query(mordred_engine* e, const mordred_star_query& q, QueryEst& out) {
  Resolved r;
  const int rc = resolve(e, &q, r);
  if (rc) return rc;
  double sel = 1.0;
  for (int k = 0; k < 4; k++) {
    if (!batch_probes(q, r, k)) continue;
    sel *= dim_pass_fraction(r, k);
    out.njoins++;
    out.dims |= 1 << k;
  }
  for (int i = 0; i < q.nfilters; i++) {
    Column* c = r.filt[i];
    column_range(c);
    const double lo = std::max<double>(q.filters[i].lo, c->vmin), hi = std::min<double>(q.filters[i].hi, c->vmax);
    sel *= hi >= lo ? (hi - lo + 1) / ((double)c->vmax - c->vmin + 1) : 0.0;
  }
  out.sel = sel;
  out.G = q.kind == MORDRED_KIND_GROUPBY ? q.total_val : 1;
  return 0;
}

// Keys of a group-by query decode from its group index when the joined
// dimensions with a non-zero stride form a mixed radix (each stride divides
// the next, the largest divides total_val) and every row's digit value -
// min_val lies in [0, next stride / stride): the group value's column range
// for a grouped dimension, 0 for a joined dimension without group value.  A
// group value with stride 0 (its key word is not in the index) does not
// decode.  Fills minv / krange for the keys kernel.
bool keys_decode(const mordred_star_query& q, const Resolved& r, mordred::BatchKeys& K) {
  int idx[4], m = 0;
  for (int k = 0; k < 4; k++) {
    K.minv[k] = 0;
    K.krange[k] = 0;
    if (!r.join[k]) continue;
    const mordred_join& j = *r.join[k];
    if (j.unique_val < 0) return false;
    if (j.unique_val == 0) {
      if (r.dval[k]) return false;
      continue;
    }
    idx[m++] = k;
  }
  std::sort(idx, idx + m, [&](int a, int b) { return r.join[a]->unique_val < r.join[b]->unique_val; });
  for (int i = 0; i < m; i++) {
    const int k = idx[i];
    const mordred_join& j = *r.join[k];
    const long long next = i + 1 < m ? r.join[idx[i + 1]]->unique_val : q.total_val;
    if (next % j.unique_val != 0) return false;
    const long long range = next / j.unique_val;
    long long lo = 0, hi = 0;
    if (r.dval[k]) {
      column_range(r.dval[k]);
      lo = r.dval[k]->vmin;
      hi = r.dval[k]->vmax;
    }
    if (lo - j.min_val < 0 || hi - j.min_val >= range) return false;
    K.minv[k] = j.min_val;
    K.krange[k] = (int)range;
  }
  return true;
}

// forms: the dimension table forms allowed (MORDRED_BATCH_FORM_*).
// lds_short is set when the plan fails because its LDS tables do not fit.
// Whether two star queries are the same query (every field; column names by
// value).
bool same_name(const char* a, const char* b) { return a == b || (a && b && std::strcmp(a, b) == 0); }
bool same_query(const mordred_star_query& a, const mordred_star_query& b) {
  if (a.njoins != b.njoins || a.nfilters != b.nfilters || a.kind != b.kind || a.total_val != b.total_val ||
      !same_name(a.aggr1, b.aggr1) || !same_name(a.aggr2, b.aggr2))
    return false;
  for (int k = 0; k < a.njoins; k++) {
    const mordred_join &x = a.joins[k], &y = b.joins[k];
    if (x.slot != y.slot || x.dim_len != y.dim_len || x.min_key != y.min_key || x.filter_mode != y.filter_mode ||
        x.c1 != y.c1 || x.c2 != y.c2 || x.min_val != y.min_val || x.unique_val != y.unique_val ||
        !same_name(x.fact_key, y.fact_key) || !same_name(x.dim_key, y.dim_key) ||
        !same_name(x.filter_col, y.filter_col) || !same_name(x.group_col, y.group_col))
      return false;
  }
  for (int f = 0; f < a.nfilters; f++)
    if (a.filters[f].lo != b.filters[f].lo || a.filters[f].hi != b.filters[f].hi ||
        !same_name(a.filters[f].col, b.filters[f].col))
      return false;
  return true;
}

int plan_batch_forms(mordred_engine* e, const mordred_star_query* const* qp, int nq, const mordred::BatchConfig& cfg,
                     BatchPlan& b, const SlotArenas* fixed, const double* sel, int forms, bool& lds_short) {
  using namespace mordred;
  lds_short = false;
  b.lds_queries = b.global_queries = b.dead = b.lds32_queries = b.keys_in_loop = 0;  // (a plan may be retried)
  if (nq < 1 || nq > kBatchMaxQ) return fail(e, "a batch holds 1..64 queries");
  if (!batch_config_ok(cfg)) return fail(e, "bad batch configuration");
  if ((size_t)cfg.lds_bytes < sizeof(BatchHot) + batch_pair_bytes(cfg.block) + kBatchMaxQ * sizeof(BatchQuery))
    return fail(e, "lds_bytes below the staged descriptor, the pair lists and the query descriptors");
  std::vector<Resolved> rs(nq);
  int first = -1;
  for (int i = 0; i < nq; i++) {
    if (!qp[i]) continue;
    const int rc = resolve(e, qp[i], rs[i]);
    if (rc) return rc;
    if (first < 0) first = i;
  }
  if (first < 0) return fail(e, "a batch needs at least one query");
  BatchDesc& D = b.d;
  std::memset(&D, 0, sizeof D);
  D.nq = nq;
  b.n = rs[first].n;
  b.res_off.assign(nq, 0);
  b.G.assign(nq, 0);
  Column* cols[kBatchMaxCols] = {};
  int ncols = 4;
  auto slot_of = [&](Column* c) -> int {
    for (int i = 0; i < ncols; i++)
      if (cols[i] == c) return i;
    if (ncols == kBatchMaxCols) return -1;
    cols[ncols] = c;
    return ncols++;
  };
  for (int i = 0; i < nq; i++) {
    if (!qp[i]) continue;
    const Resolved& r = rs[i];
    if (r.n != b.n) return fail(e, "batch queries read fact tables of different lengths");
    for (Column* c : r.fact_cols)
      for (int s = 0; s < c->nseg; s++)
        if (c->slot[s] < 0) return fail(e, "batch fact columns must be cached in HBM: " + c->name);
    for (int k = 0; k < 4; k++) {
      if (!batch_probes(*qp[i], r, k)) continue;
      if (cols[k] && cols[k] != r.key[k]) return fail(e, "batch queries join one table through different fact keys");
      cols[k] = r.key[k];
    }
  }
  // dimensions: mask form (a query-mask table whose 16-byte entries also
  // hold the group attributes when a query groups on the dimension) when the
  // mask table stays small,
  // else attribute form (one table of every filtered or grouped attribute);
  // key filters of attribute-form dimensions become constraints on the fact
  // key column's value (kcons), evaluated by its interval table
  struct Cons {
    int q;
    int nr;
    long long lo[2], hi[2];  // accepted values: the union of [lo, hi] ranges
  };
  std::vector<Cons> kcons[4];
  auto filter_ranges = [](int mode, int c1, int c2, Cons& c) {
    if (mode == 0) { c.nr = 1; c.lo[0] = c.hi[0] = c1; }
    else if (mode == 1) { c.nr = 1; c.lo[0] = c1; c.hi[0] = c2; }
    else { c.nr = 2; c.lo[0] = c.hi[0] = c1; c.lo[1] = c.hi[1] = c2; }
  };
  Column* attrs[4][kBatchMaxAttr] = {};
  std::vector<char> qdead(nq, 0);  // a dimension filter with an empty range
  D.ntab = 0;
  D.attr16 = 1;
  D.uniq16 = 1;
  for (int k = 0; k < 4; k++) {
    BatchDim& B = D.dim[k];
    Column* key = nullptr;
    struct DFilt {
      int q, mode, c1, c2;
      Column* col;
    };
    std::vector<DFilt> filts;
    std::vector<Column*> gcols;
    for (int i = 0; i < nq; i++) {
      if (!qp[i]) continue;
      const Resolved& r = rs[i];
      if (!batch_probes(*qp[i], r, k)) continue;
      const mordred_join& j = *r.join[k];
      for (Column* c : {r.dkey[k], r.dfilt[k], r.dval[k]}) {
        if (c && !c->pinned_gpu) return fail(e, "dimension columns must be GPU-resident");
        if (c && c->n < r.dkey[k]->n) return fail(e, "dimension column shorter than its key column: " + c->name);
      }
      if (!key) {
        key = r.dkey[k];
        B.key = key->d_idx;
        B.len = j.dim_len;
        B.minkey = j.min_key;
        B.nrows = (int)key->n;
      } else if (key != r.dkey[k] || B.len != j.dim_len || B.minkey != j.min_key) {
        return fail(e, "batch queries disagree on a dimension's key, dim_len or min_key");
      }
      B.uses |= 1ull << i;
      if (r.dfilt[k]) {
        if (j.filter_mode < 0 || j.filter_mode > 2) return fail(e, "dimension filter mode must be 0..2");
        // mode 1 with c1 > c2 keeps no dimension row: the query is dead
        if (j.filter_mode == 1 && j.c1 > j.c2) qdead[i] = 1;
        filts.push_back(DFilt{i, j.filter_mode, j.c1, j.c2, r.dfilt[k]});
      }
      if (qp[i]->kind == MORDRED_KIND_GROUPBY && r.dval[k]) {
        B.groupers |= 1ull << i;
        if (std::find(gcols.begin(), gcols.end(), r.dval[k]) == gcols.end()) gcols.push_back(r.dval[k]);
      }
    }
    if (!key) continue;
    // no two keys may share a slot: the dimension's and the fact's key ranges
    // lie within dim_len, so the row in a probed slot has the fact row's key
    Column* fkey = cols[k];
    column_range(key);
    column_range(fkey);
    const long long klo = std::min(key->vmin, fkey->vmin), khi = std::max(key->vmax, fkey->vmax);
    if (khi - klo >= (long long)B.len)
      return fail(e, "dimension key range exceeds dim_len (hash collisions possible): batches need perfect hashing");
    // form: attribute form for a large table whose non-key filters are all on
    // attributes narrow enough for a mask table per value
    std::vector<Column*> fattrs;
    bool narrow = true;
    for (const DFilt& f : filts) {
      if (f.col == key) continue;
      column_range(f.col);
      if ((long long)f.col->vmax - f.col->vmin + 1 > kBatchLutSpan) narrow = false;
      if (std::find(fattrs.begin(), fattrs.end(), f.col) == fattrs.end()) fattrs.push_back(f.col);
    }
    std::vector<Column*> all_attrs = gcols;
    for (Column* c : fattrs)
      if (std::find(all_attrs.begin(), all_attrs.end(), c) == all_attrs.end()) all_attrs.push_back(c);
    // dictionary form when the rows hold few distinct attribute tuples
    const DimDict* dict = nullptr;
    if ((forms & MORDRED_BATCH_FORM_DICT) && (int)all_attrs.size() <= kBatchMaxAttr) {
      const int rc = dim_dict(e, key, all_attrs, B.len, B.minkey, &dict);
      if (rc) return rc;
      if (dict && ((D.ntab + 3) & ~3) + 4 * ((int)dict->tuples.size() + 1) > kBatchMaxTab) dict = nullptr;
    }
    B.form = dict ? kDimDict
                  : ((forms & MORDRED_BATCH_FORM_ATTR) &&
                     ((long long)B.len * 8 > kBatchMaskFormBytes || !(forms & MORDRED_BATCH_FORM_MASK)) &&
                     narrow && (int)all_attrs.size() <= kBatchMaxAttr)
                        ? kDimAttr : kDimMask;
    const std::vector<Column*>& packed = B.form == kDimDict ? dict->attrs : B.form == kDimAttr ? all_attrs : gcols;
    if ((int)packed.size() > kBatchMaxAttr) return fail(e, "more than 3 group columns on one dimension in a batch");
    // packed attributes from bit 0 (attribute form: value - base + 1, so one
    // more value than the span)
    int off = 0;
    for (size_t a = 0; a < packed.size(); a++) {
      Column* c = packed[a];
      attrs[k][a] = c;
      column_range(c);
      const long long span = (long long)c->vmax - c->vmin + 1 + (B.form != kDimMask ? 1 : 0);
      int bits = 1;
      while ((1ll << bits) < span) bits++;
      B.attr[a] = BatchAttr{c->d_idx, c->vmin, off, bits, -1, 0};
      off += bits;
    }
    if (off > 32) return fail(e, "attributes of one dimension exceed 32 bits in a batch");
    if (off > 16) D.attr16 = 0;
    B.nattr = (int)packed.size();
    B.pw = off <= 8 ? 1 : off <= 16 ? 2 : 4;
    if (B.form == kDimMask) {
      // filters evaluated by the table build, per dimension row
      Column* fcols[kBatchMaxFCol] = {};
      int nfc = 0;
      for (const DFilt& f : filts) {
        int src = -1;
        if (f.col != key) {
          src = 0;
          while (src < nfc && fcols[src] != f.col) src++;
          if (src == nfc) {
            if (nfc == kBatchMaxFCol) return fail(e, "more than 4 filtered columns on one dimension in a batch");
            fcols[nfc++] = f.col;
          }
        }
        if (B.nfilt == kBatchMaxFilt) return fail(e, "too many dimension filters in a batch");
        B.filt[B.nfilt++] = BatchFilt{src, f.mode, f.c1, f.c2, f.q, {0, 0, 0}};
      }
      B.nfcol = nfc;
      for (int c = 0; c < nfc; c++) B.fcol[c] = fcols[c]->d_idx;
      // slot len stays zero: the probe of an item no query needs, or whose
      // key has no slot, reads it
      B.mstride = B.nattr ? 2 : 1;
      const size_t mbytes = (((size_t)(B.len + 1) * 8 * B.mstride) + 15) & ~(size_t)15;
      // (growing frees the buffer: the request loop may plan while a launch
      // that reads it is in flight)
      if (mbytes > e->d_mk[k].bytes) HIP_OK(hipStreamSynchronize(e->s_main));
      HIP_OK(e->d_mk[k].reserve(mbytes));
      B.mk = static_cast<unsigned long long*>(e->d_mk[k].p);
    } else if (B.form == kDimDict) {
      // per code 16 bytes: the mask of the queries whose attribute filters
      // its tuple passes, its packed attributes, 0 (code 0: empty slot)
      const int nt = (int)dict->tuples.size() + 1;
      const int at = (D.ntab + 3) & ~3;
      B.dict_off = at;
      B.cb = dict->cb;
      for (int t = 0; t < nt; t++) {
        const unsigned w = t ? dict->tuples[t - 1] : 0u;
        unsigned long long m = t ? ~0ull : 0ull;
        for (const DFilt& f : filts) {
          if (f.col == key || !t) continue;
          int a = 0;
          while (attrs[k][a] != f.col) a++;
          const unsigned fm = B.attr[a].bits >= 32 ? ~0u : (1u << B.attr[a].bits) - 1u;
          const int x = (int)((w >> B.attr[a].off) & fm) - 1 + B.attr[a].base;
          const bool ok = f.mode == 0 ? x == f.c1 : f.mode == 1 ? (x >= f.c1 && x <= f.c2) : (x == f.c1 || x == f.c2);
          if (!ok) m &= ~(1ull << f.q);
        }
        D.tab[at + 4 * t] = (unsigned)m;
        D.tab[at + 4 * t + 1] = (unsigned)(m >> 32);
        D.tab[at + 4 * t + 2] = w;
        D.tab[at + 4 * t + 3] = 0u;
      }
      D.ntab = at + 4 * nt;
      for (const DFilt& f : filts) {
        if (f.col != key) continue;
        Cons c{f.q, 0, {0, 0}, {0, 0}};
        filter_ranges(f.mode, f.c1, f.c2, c);
        kcons[k].push_back(c);
      }
      B.mk = nullptr;
      B.nfilt = 0;
      B.nfcol = 0;
      B.pk = dict->codes.p;
    } else {
      // a mask table per filtered attribute (bit q cleared for the values q
      // rejects); key filters go to the fact key column's interval table
      for (int a = 0; a < B.nattr; a++) {
        Column* c = attrs[k][a];
        if (std::find(fattrs.begin(), fattrs.end(), c) == fattrs.end()) continue;
        const int span = (int)((long long)c->vmax - c->vmin + 1);
        const int at = (D.ntab + 1) & ~1;
        if (at + 2 * span > kBatchMaxTab) {
          lds_short = true;
          return fail(e, "attribute mask tables exceed the batch's LDS table space");
        }
        B.attr[a].lut = at;
        for (int v = 0; v < span; v++) {
          unsigned long long m = ~0ull;
          for (const DFilt& f : filts) {
            if (f.col != c) continue;
            const int x = c->vmin + v;
            const bool ok = f.mode == 0 ? x == f.c1 : f.mode == 1 ? (x >= f.c1 && x <= f.c2) : (x == f.c1 || x == f.c2);
            if (!ok) m &= ~(1ull << f.q);
          }
          D.tab[at + 2 * v] = (unsigned)m;
          D.tab[at + 2 * v + 1] = (unsigned)(m >> 32);
        }
        D.ntab = at + 2 * span;
      }
      for (const DFilt& f : filts) {
        if (f.col != key) continue;
        Cons c{f.q, 0, {0, 0}, {0, 0}};
        filter_ranges(f.mode, f.c1, f.c2, c);
        kcons[k].push_back(c);
      }
      B.mk = nullptr;
      B.nfilt = 0;
      B.nfcol = 0;
    }
    B.used = 1;
    if (B.form == kDimAttr) {
      // a zero entry at slot len, and the aligned 8-byte word holding it
      const size_t pbytes = (((size_t)(B.len + 1) * B.pw + 8) + 15) & ~(size_t)15;
      if (pbytes > e->d_rec[k].bytes) HIP_OK(hipStreamSynchronize(e->s_main));
      HIP_OK(e->d_rec[k].reserve(pbytes));
      B.pk = e->d_rec[k].p;
      B.pk_words = (long long)(pbytes / 16);
    } else if (B.form == kDimDict) {
      B.pk_words = 0;  // (the cached code table, never cleared)
    } else {
      B.pk = nullptr;
      B.pk_words = 0;
    }
  }
  // fact predicates, aggregate expressions, query descriptors
  std::vector<BatchPred> preds;
  for (int i = 0; i < nq; i++) {
    if (!qp[i]) continue;
    const mordred_star_query& q = *qp[i];
    const Resolved& r = rs[i];
    bool dead = qdead[i] != 0;
    for (int f = 0; f < q.nfilters; f++) {
      const int sl = slot_of(r.filt[f]);
      if (sl < 0) return fail(e, "a batch reads more than 12 fact columns");
      if (q.filters[f].lo > q.filters[f].hi) dead = true;
      preds.push_back(BatchPred{sl, q.filters[f].lo, q.filters[f].hi, i});
    }
    const int kind = q.kind == MORDRED_KIND_GROUPBY ? kExprDiff : kExprProduct;
    const int a1 = slot_of(r.a1), a2 = r.a2 ? slot_of(r.a2) : -1;
    if (a1 < 0 || (r.a2 && a2 < 0)) return fail(e, "a batch reads more than 12 fact columns");
    int ex = -1;
    for (int x = 0; x < D.nexpr; x++)
      if (D.expr[x].kind == kind && D.expr[x].a1 == a1 && D.expr[x].a2 == a2) ex = x;
    if (ex < 0) {
      if (D.nexpr == kBatchMaxExpr) return fail(e, "more than 4 distinct aggregate expressions in a batch");
      ex = D.nexpr++;
      D.expr[ex] = BatchExpr{kind, a1, a2, 0};
    }
    BatchQuery& Q = D.q[i];
    Q.total_val = q.kind == MORDRED_KIND_GROUPBY ? q.total_val : 1;
    Q.expr = (unsigned char)ex;
    Q.flags = (unsigned char)(q.kind == MORDRED_KIND_GROUPBY ? 0 : kQueryAggr);
    Q.gtouch = 0;
    unsigned hbase = 0;
    for (int k = 0; k < 4; k++) {
      Q.attr[k] = -1;
      Q.shift[k] = 0;
      Q.bits[k] = 0;
      Q.uniq[k] = 0;
      if (q.kind != MORDRED_KIND_GROUPBY || !r.join[k]) continue;
      const int uv = r.join[k]->unique_val;
      if (uv < 0 || uv >= (1 << 24)) return fail(e, "batch group index needs unique_val in [0, 2^24)");
      Q.uniq[k] = uv;
      if (uv >= (1 << 16)) D.uniq16 = 0;
      if (!r.dval[k]) {
        // a joined table without group value contributes its probe's value
        // word, 0 (GPUProcessing.h:142): (0 - min_val) * unique_val
        hbase += (0u - (unsigned)r.join[k]->min_val) * (unsigned)uv;
        continue;
      }
      int a = 0;
      while (a < kBatchMaxAttr && attrs[k][a] != r.dval[k]) a++;
      const BatchAttr& A = D.dim[k].attr[a];
      Q.attr[k] = (signed char)a;
      // (16-bit attribute words: dimensions 1 and 3 ride in the high halves of
      // 0 and 2 during the scan's pair loop)
      Q.shift[k] = (unsigned char)(A.off + (D.attr16 && (k & 1) ? 16 : 0));
      Q.bits[k] = (unsigned char)A.bits;
      if (A.bits > 16) return fail(e, "batch group index needs attribute fields of at most 16 bits");
      // the field holds value - base (+ 1 in the attribute and dictionary forms)
      const unsigned fbase = (unsigned)A.base - (D.dim[k].form != kDimMask ? 1u : 0u);
      hbase += (fbase - (unsigned)r.join[k]->min_val) * (unsigned)uv;
    }
    Q.hbase = (int)hbase;
    if (q.kind == MORDRED_KIND_GROUPBY && !keys_decode(q, r, D.keys[i])) Q.flags |= kQueryKeysInLoop;
    b.G[i] = Q.total_val;
    if (dead) b.dead++;
    else D.all |= 1ull << i;
  }
  std::stable_sort(preds.begin(), preds.end(), [](const BatchPred& x, const BatchPred& y) { return x.col < y.col; });
  // interval tables: a fact column with at least kBatchIvlMin predicates, or
  // the fact key of an attribute-form dimension with key filters; their
  // predicates leave the per-predicate compare list
  D.nivl = 0;
  for (int c = 0; c < kBatchMaxCols; c++) {
    std::vector<Cons> on;
    int npc = 0;
    for (const BatchPred& p : preds)
      if (p.col == c) {
        Cons x{p.q, 1, {p.lo, 0}, {p.hi, 0}};
        on.push_back(x);
        npc++;
      }
    const bool keyed = c < 4 && !kcons[c].empty();
    if (keyed) on.insert(on.end(), kcons[c].begin(), kcons[c].end());
    if (!keyed && npc < kBatchIvlMin) continue;
    if (D.nivl == kBatchMaxIvl) {
      if (keyed) {
        lds_short = true;
        return fail(e, "too many interval tables in a batch");
      }
      continue;
    }
    column_range(cols[c]);
    const long long vmin = cols[c]->vmin, vmax = cols[c]->vmax;
    // breakpoints that split the column's value range [vmin, vmax]: the
    // range ends lo and hi + 1 inside (vmin, vmax]
    std::vector<long long> pts;
    for (const Cons& x : on)
      for (int r = 0; r < x.nr; r++)
        for (long long v : {x.lo[r], x.hi[r] + 1})
          if (v > vmin && v <= vmax) pts.push_back(v);
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    const int n = (int)pts.size();
    int shift = 0;
    while (((vmax - vmin) >> shift) + 1 > kBatchMaxBuckets) shift++;
    const int nb = (int)(((vmax - vmin) >> shift) + 1);
    // rank at each bucket start, and the most breakpoints inside one bucket
    std::vector<int> rank(nb);
    int kmax = 0;
    for (int t = 0, i = 0; t < nb; t++) {
      const long long start = vmin + ((long long)t << shift), stop = start + (1ll << shift);
      while (i < n && pts[i] <= start) i++;  // breakpoints at or below the start count for every value here
      rank[t] = i;
      int in = 0;
      while (i + in < n && pts[i + in] < stop) in++;
      kmax = std::max(kmax, in);
    }
    const int mask_words = 2 * (n + 1), rank_words = (nb + 1) / 2, bp_words = n + kmax;
    const int at = (D.ntab + 1) & ~1;
    if (at + mask_words + rank_words + bp_words > kBatchMaxTab) {
      if (keyed) {
        lds_short = true;
        return fail(e, "interval tables exceed the batch's LDS table space");
      }
      continue;
    }
    BatchIvl& I = D.ivl[D.nivl++];
    I.col = c;
    I.base = (int)vmin;
    I.shift = shift;
    I.nb = nb;
    I.n = n;
    I.kmax = kmax;
    I.mask_off = at;
    I.rank_off = I.mask_off + mask_words;
    I.bp_off = I.rank_off + rank_words;
    D.ntab = I.bp_off + bp_words;
    // interval r = [pts[r-1], pts[r]), the first from vmin, the last to vmax:
    // a query accepts it iff it lies inside one of the ranges of each of the
    // query's constraints on the column
    for (int r = 0; r <= n; r++) {
      const long long lo_r = r == 0 ? vmin : pts[r - 1];
      const long long hi_r = r == n ? vmax + 1 : pts[r];  // exclusive
      unsigned long long m = ~0ull;
      for (const Cons& x : on) {
        bool in = false;
        for (int k2 = 0; k2 < x.nr; k2++) in |= lo_r >= x.lo[k2] && hi_r <= x.hi[k2] + 1;
        if (!in) m &= ~(1ull << x.q);
      }
      D.tab[I.mask_off + 2 * r] = (unsigned)m;
      D.tab[I.mask_off + 2 * r + 1] = (unsigned)(m >> 32);
    }
    for (int t = 0; t < rank_words; t++) {
      const unsigned lo16 = (unsigned)rank[2 * t], hi16 = 2 * t + 1 < nb ? (unsigned)rank[2 * t + 1] : 0u;
      D.tab[I.rank_off + t] = lo16 | hi16 << 16;
    }
    for (int i = 0; i < bp_words; i++) D.tab[I.bp_off + i] = i < n ? (unsigned)(int)pts[i] : (unsigned)INT_MAX;
    preds.erase(std::remove_if(preds.begin(), preds.end(), [&](const BatchPred& p) { return p.col == c; }),
                preds.end());
  }
  D.npred = (int)preds.size();
  for (int p = 0; p < D.npred; p++) D.pred[p] = preds[p];
  for (int c = 0, p = 0; c <= kBatchMaxCols; c++) {
    while (p < D.npred && D.pred[p].col < c) p++;
    D.pred_begin[c] = p;
  }
  D.ncols = ncols;
  b.cols_read = 0;
  for (int c = 0; c < ncols; c++) {
    D.col[c] = cols[c] ? cols[c]->d_idx : nullptr;
    b.cols_read += cols[c] != nullptr;
  }
  // group sinks.  Every query with decodable keys keeps a touched bitmap in
  // LDS; the rest of the LDS budget holds accumulators, given in the order of
  // expected pairs per LDS word (selectivity / groups): first as 32-bit
  // accumulators, then upgraded to 64-bit ones (no returning atomic) while
  // room is left.  The others add per pair to the result (global atomics).
  std::vector<double> est_sel(nq, 0.0);
  for (int i = 0; i < nq; i++) {
    if (!qp[i] || !(D.all >> i & 1ull)) continue;
    if (sel) {
      est_sel[i] = sel[i];
    } else {
      QueryEst qe;
      if (estimate_query(e, *qp[i], qe) == 0) est_sel[i] = qe.sel;
    }
  }
  const long long fixed_bytes = (long long)sizeof(BatchHot) + batch_pair_bytes(cfg.block) +
                                kBatchMaxQ * (long long)sizeof(BatchQuery) + 4LL * ((D.ntab + 1) & ~1);
  if (fixed_bytes > cfg.lds_bytes) {
    lds_short = true;
    return fail(e, "lds_bytes below the staged descriptor and interval tables");
  }
  long long avail = ((long long)cfg.lds_bytes - fixed_bytes) / 4;  // 32-bit words
  std::vector<int> order;
  for (int i = 0; i < nq; i++)
    if (qp[i]) order.push_back(i);
  // Order: the time a query would add on global atomics per LDS word it
  // takes -- its pairs (selectivity) per group, weighted up for few groups,
  // whose atomics all land on a few addresses and serialize there.  An
  // estimate below 1e-3 counts as 1e-3 (estimates of small selectivities are
  // the least reliable, and a small group array costs little LDS).
  // identical live queries (a request stream repeats queries): the first
  // is the primary; the others can share its LDS sums (below), so the
  // primary's sums are worth as much LDS as all of theirs
  std::vector<int> primary(nq, -1), copies(nq, 1);
  if (e->batch_share)
    for (int i = 0; i < nq; i++) {
      if (!qp[i] || !(D.all >> i & 1ull) || (D.q[i].flags & kQueryKeysInLoop)) continue;
      for (int j = 0; j < i; j++)
        if (primary[j] < 0 && qp[j] && (D.all >> j & 1ull) && !(D.q[j].flags & kQueryKeysInLoop) &&
            same_query(*qp[i], *qp[j])) {
          primary[i] = j;
          copies[j]++;
          break;
        }
    }
  std::vector<double> prio(nq, 0.0);
  for (int i = 0; i < nq; i++) {
    const double g = std::max(1, b.G[i]);
    prio[i] = primary[i] >= 0 ? 0.0 : copies[i] * std::max(est_sel[i], 1e-3) / g * (1.0 + 64.0 / g);
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return prio[x] > prio[y]; });
  // touched bitmaps first; if they do not fit, the largest queries store
  // their keys in the pair loop instead (no bitmap)
  {
    long long need = 0;
    for (int i : order) need += batch_touch_words(b.G[i], D.q[i].flags);
    std::vector<int> by_g = order;
    std::stable_sort(by_g.begin(), by_g.end(), [&](int x, int y) { return b.G[x] > b.G[y]; });
    for (int i : by_g) {
      if (need <= avail) break;
      const int tw = batch_touch_words(b.G[i], D.q[i].flags);
      if (!tw) continue;
      D.q[i].flags |= kQueryKeysInLoop;
      need -= tw;
    }
    avail -= need;
  }
  std::vector<int> sink(nq, kSinkGlobal);
  auto even = [](long long w) { return (w + 1) & ~1ll; };
  for (int i : order) {
    if (!(D.all >> i & 1ull)) continue;
    // queries that store keys in the pair loop need the 64-bit sum back
    const int s0 = (D.q[i].flags & kQueryKeysInLoop) ? kSinkLds64 : kSinkLds32;
    const long long w = even(batch_acc_words(s0, b.G[i]));
    if (w <= avail) {
      sink[i] = s0;
      avail -= w;
    }
  }
  // identical queries whose primary got LDS accumulators take its sums
  // instead of scanning: the rows of one launch count alike for both
  D.aliases = 0;
  for (int i = 0; i < nq; i++) D.q[i].alias = -1;
  for (int i = 0; i < nq; i++) {
    const int p = primary[i];
    if (p < 0 || sink[p] == kSinkGlobal || ((D.q[p].flags | D.q[i].flags) & kQueryKeysInLoop)) continue;
    D.all &= ~(1ull << i);
    D.aliases |= 1ull << i;
    sink[i] = kSinkGlobal;  // (no accumulators of its own)
    // append to the primary's chain
    int t = p;
    while (D.q[t].alias >= 0) t = D.q[t].alias;
    D.q[t].alias = i;
    b.shared++;
  }
  for (int i : order) {
    if (sink[i] != kSinkLds32) continue;
    const long long extra = even(batch_acc_words(kSinkLds64, b.G[i])) - even(batch_acc_words(kSinkLds32, b.G[i]));
    if (extra <= avail) {
      sink[i] = kSinkLds64;
      avail -= extra;
    }
  }
  // the group region follows the query descriptors and the tables in the
  // scan's dynamic LDS; per query its touched bitmap, then its accumulators
  const long long grp0 = kBatchMaxQ * (long long)sizeof(BatchQuery) / 4 + ((D.ntab + 1) & ~1);
  long long lds_words = 0, res_ints = 0, touch_words = 0;
  for (int i = 0; i < nq; i++) {
    BatchQuery& Q = D.q[i];
    Q.sink = (unsigned char)sink[i];
    Q.touch = kBatchNoTouch;
    Q.acc = (int)(4 * (grp0 + lds_words));
    if (!qp[i]) continue;
    const int tw = batch_touch_words(b.G[i], Q.flags);
    if (tw) Q.touch = (unsigned short)(grp0 + lds_words);
    Q.acc = (int)(4 * (grp0 + lds_words + tw));
    lds_words += tw + even(batch_acc_words(sink[i], b.G[i]));
    if (D.aliases >> i & 1ull) {
      // (its touched bitmap and accumulators stay unused: the primary's serve it)
    } else if (sink[i] == kSinkGlobal) {
      b.global_queries++;
    } else {
      b.lds_queries++;
    }
    b.lds32_queries += sink[i] == kSinkLds32;
    b.keys_in_loop += (Q.flags & kQueryKeysInLoop) != 0;
    if (fixed && 6 * (long long)b.G[i] > fixed->res_stride)
      return fail(e, "query result larger than its serving slot");
    if (fixed && tw > fixed->touch_stride) return fail(e, "query groups exceed its serving slot's touched bitmap");
    b.res_off[i] = fixed ? i * fixed->res_stride : res_ints;
    Q.gtouch = fixed ? i * fixed->touch_stride : (int)touch_words;
    res_ints += 6 * (long long)b.G[i];
    touch_words += tw;
  }
  D.lds_words = (int)lds_words;
  D.lds_dyn = cfg.lds_bytes - (int)sizeof(BatchHot);
  if (4 * (grp0 + lds_words) + batch_pair_bytes(cfg.block) > D.lds_dyn || grp0 + lds_words >= kBatchNoTouch)
    return fail(e, "batch LDS plan exceeds the launch's LDS");
  if (fixed) {
    D.touch_words = nq * fixed->touch_stride;
    D.touch_arena = fixed->touch;
  } else {
    D.touch_words = (int)touch_words;
    HIP_OK(e->d_touch.reserve(std::max<long long>(16, touch_words * 4)));
    D.touch_arena = static_cast<unsigned*>(e->d_touch.p);
  }
  if (fixed) {
    D.res_ints = nq * fixed->res_stride;
    D.res_arena = fixed->res;
  } else {
    D.res_ints = (res_ints + 3) & ~3ll;
    HIP_OK(e->d_arena.reserve(std::max<long long>(16, D.res_ints * 4)));
    D.res_arena = static_cast<int*>(e->d_arena.p);
  }
  for (int i = 0; i < nq; i++) {
    D.q[i].res = (long long)(uintptr_t)(D.res_arena + b.res_off[i]);
    if (qp[i]) D.reset |= 1ull << i;
  }
  // one launch over every row, record tables built
  D.seg_begin = 0;
  D.num_tuples = (int)b.n;
  D.rebuild = 1;
  return 0;
}

// A plan with the engine's allowed table forms, else (when its LDS tables do
// not fit the configuration) with the mask form only, whose tables take the
// least LDS.
int plan_batch(mordred_engine* e, const mordred_star_query* const* qp, int nq, const mordred::BatchConfig& cfg,
               BatchPlan& b, const SlotArenas* fixed = nullptr, const double* sel = nullptr) {
  bool short_ = false;
  const int rc = plan_batch_forms(e, qp, nq, cfg, b, fixed, sel, e->batch_forms, short_);
  if (rc == 0 || !short_ || e->batch_forms == MORDRED_BATCH_FORM_MASK) return rc;
  return plan_batch_forms(e, qp, nq, cfg, b, fixed, sel, MORDRED_BATCH_FORM_MASK, short_);
}

// Launch arguments and record tables a graph was captured with.
struct GraphKey {
  mordred::BatchLaunch L;
  mordred::BatchConfig c;
  int nrows[4];     // dimension rows (table build grids)
  int used[4];
  int ncols;        // selects the scan kernel's register layout
};

// Capture (first use / changed launch) and launch the batch graph of plan b on
// the main stream, bracketed by the engine's events.  Does not wait.
int launch_batch(mordred_engine* e, const BatchPlan& b, const mordred::BatchConfig& cfg, int* captured_out) {
  using namespace mordred;
  HIP_OK(hipSetDevice(e->device));
  if (!e->ev0) HIP_OK(hipEventCreate(&e->ev0));
  if (!e->ev1) HIP_OK(hipEventCreate(&e->ev1));
  HIP_OK(e->d_desc.reserve(sizeof(BatchDesc)));
  HIP_OK(e->h_desc.reserve(sizeof(BatchDesc)));
  HIP_OK(e->d_status.reserve(64));
  GraphKey k;
  std::memset(&k, 0, sizeof k);
  k.L = BatchLaunch{static_cast<const BatchDesc*>(e->d_desc.p), e->cache, e->d_identity, (int)b.n,
                    static_cast<unsigned*>(e->d_status.p)};
  k.c = cfg;
  for (int d = 0; d < 4; d++) {
    k.used[d] = b.d.dim[d].used;
    k.nrows[d] = k.used[d] ? b.d.dim[d].nrows : 0;
  }
  k.ncols = b.d.ncols <= 8 ? 8 : 12;
  int captured = 0;
  const char* kp = reinterpret_cast<const char*>(&k);
  if (!e->gexec || e->gkey.size() != sizeof k || std::memcmp(e->gkey.data(), kp, sizeof k) != 0) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
    e->gkey.clear();
    HIP_OK(batch_prepare(cfg, b.d.ncols));
    (void)batch_resident(cfg, b.d.ncols);  // occupancy query outside the capture
    HIP_OK(hipStreamBeginCapture(e->s_main, hipStreamCaptureModeRelaxed));
    hipError_t r1 = batch_fill(k.L, e->s_main);
    for (int d = 0; d < 4 && r1 == hipSuccess; d++)
      if (k.used[d]) r1 = batch_records(k.L, d, k.nrows[d], e->s_main);
    if (r1 == hipSuccess) r1 = batch_scan(k.L, cfg, b.d.ncols, e->s_main);
    if (r1 == hipSuccess) r1 = batch_keys(k.L, e->s_main);
    hipGraph_t g = nullptr;
    const hipError_t r2 = hipStreamEndCapture(e->s_main, &g);
    if (r1 != hipSuccess || r2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(e, r1 != hipSuccess ? r1 : r2, "batch graph capture");
    }
    e->graph = g;
    HIP_OK(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
    e->gkey.assign(kp, kp + sizeof k);
    captured = 1;
  }
  std::memcpy(e->h_desc.p, &b.d, sizeof(BatchDesc));
  HIP_OK(hipMemcpyAsync(e->d_desc.p, e->h_desc.p, sizeof(BatchDesc), hipMemcpyHostToDevice, e->s_main));
  HIP_OK(hipEventRecord(e->ev0, e->s_main));
  HIP_OK(hipGraphLaunch(e->gexec, e->s_main));
  HIP_OK(hipEventRecord(e->ev1, e->s_main));
  if (captured_out) *captured_out = captured;
  return 0;
}

int batch_status(mordred_engine* e, unsigned status) {
  if (status) return fail(e, "batch kernels reported status " + std::to_string(status));
  return 0;
}

int run_batch(mordred_engine* e, const BatchPlan& b, const mordred::BatchConfig& cfg, int* const* res,
              mordred_batch_stats* st) {
  using namespace mordred;
  const auto t0 = Clock::now();
  int captured = 0;
  int rc = launch_batch(e, b, cfg, &captured);
  if (rc) return rc;
  const size_t rbytes = (size_t)b.d.res_ints * sizeof(int);
  HIP_OK(e->h_arena.reserve(rbytes + 64));
  char* ha = static_cast<char*>(e->h_arena.p);
  if (rbytes) HIP_OK(hipMemcpyAsync(ha, b.d.res_arena, rbytes, hipMemcpyDeviceToHost, e->s_main));
  HIP_OK(hipMemcpyAsync(ha + rbytes, e->d_status.p, sizeof(unsigned), hipMemcpyDeviceToHost, e->s_main));
  HIP_OK(hipStreamSynchronize(e->s_main));
  unsigned collisions = 0;
  std::memcpy(&collisions, ha + rbytes, sizeof collisions);
  if ((rc = batch_status(e, collisions))) return rc;
  const int* arena = reinterpret_cast<const int*>(ha);
  for (int i = 0; i < b.d.nq; i++)
    std::memcpy(res[i], arena + b.res_off[i], 6 * (size_t)b.G[i] * sizeof(int));
  if (st) {
    float gms = 0;
    HIP_OK(hipEventElapsedTime(&gms, e->ev0, e->ev1));
    st->gpu_ms = gms;
    st->total_ms = ms_since(t0);
    st->lds_queries = b.lds_queries;
    st->global_queries = b.global_queries;
    st->dead_queries = b.dead;
    st->graph_captured = captured;
    st->config = mordred_batch_config{cfg.block, cfg.vpt, cfg.lds_bytes};
    st->alg_bytes = 4.0 * b.cols_read * (double)b.d.num_tuples;
    st->lds32_queries = b.lds32_queries;
    st->keys_in_loop_queries = b.keys_in_loop;
    st->shared_queries = b.shared;
  }
  return 0;
}

// ------------------------------------------------------------ request loop
// Selectivity estimate of one query (the fraction of fact rows it keeps):
// dimension filters on an evenly spaced sample of the dimension rows, fact
// BETWEEN filters from the column's value range assuming uniform values.
// Features of one launch of the shared scan (MORDRED_COST_FEATURES floats).
void launch_features(const BatchPlan& b, const std::vector<QueryEst>& est, long long rows, int nreset,
                     const mordred::BatchConfig& c, float* f) {
  std::fill(f, f + MORDRED_COST_FEATURES, 0.f);
  const mordred::BatchDesc& D = b.d;
  double s_all = 0, s_glob = 0, s_lds = 0, g_all = 0, g_lds = 0, gmax = 0;
  int nq = 0, nglob = 0, deep = 0;
  int per_dim[4] = {0, 0, 0, 0};
  for (int i = 0; i < D.nq; i++) {
    if (!(D.all >> i & 1ull)) continue;
    const QueryEst& q = est[i];
    nq++;
    s_all += q.sel;
    g_all += q.G;
    gmax = std::max<double>(gmax, q.G);
    if (D.q[i].sink != mordred::kSinkGlobal) {
      s_lds += q.sel;
      g_lds += q.G;
    } else {
      s_glob += q.sel;
      nglob++;
    }
    deep += q.njoins >= 3;
    for (int k = 0; k < 4; k++) per_dim[k] += q.dims >> k & 1;
  }
  f[0] = (float)(std::log2((double)rows + 1.0) / 32.0);
  f[1] = nq / 64.f;
  f[2] = (float)s_all;
  f[3] = (float)s_glob;
  f[4] = (float)s_lds;
  f[5] = (float)(std::log2(1.0 + g_all) / 16.0);
  f[6] = nglob / 64.f;
  f[7] = D.ncols / 12.f;
  for (int k = 0; k < 4; k++) {
    f[8 + k] = (float)D.dim[k].used;
    f[12 + k] = per_dim[k] / 64.f;
  }
  f[16] = D.npred / 128.f;
  f[17] = D.nexpr / 4.f;
  f[18] = c.block == 256;
  f[19] = c.block == 512;
  f[20] = c.block == 1024;
  f[21] = (float)(c.vpt - 1);
  f[22] = c.lds_bytes / (160.f * 1024.f);
  f[23] = (float)D.rebuild;
  f[24] = nreset / 64.f;
  f[25] = (float)(std::log2(1.0 + g_lds) / 16.0);
  f[26] = (float)(gmax / 17500.0);
  f[27] = nq ? (float)deep / nq : 0.f;
  f[28] = (float)(rows * s_all / 1e9);
  f[29] = (float)(rows * s_glob / 1e9);
  f[30] = (float)(rows / 1e9);
  f[31] = 1.f;
}

// Analytic launch model: column bytes at the shared scan's streaming rate,
// one LDS or global update per surviving (row, query) pair, a fixed launch
// cost and the record rebuild.  Constants from MI355X measurements of the
// shared scan (64-query batches over 600M rows).
double analytic_launch_ms(const float* f, int ncols) {
  const double rows = f[30] * 1e9;
  const double bytes = rows * 4.0 * ncols;
  const double pairs_lds = (double)f[28] * 1e9 - (double)f[29] * 1e9;
  const double pairs_glob = (double)f[29] * 1e9;
  return 1e3 * (bytes / 1.3e12 + pairs_lds * 8e-12 + pairs_glob * 60e-12) + 0.08 + (f[23] > 0 ? 0.3 : 0.0);
}

}  // namespace

// ================================================================= C ABI
extern "C" {

mordred_engine* mordred_engine_create(int device, int cache_segments, int threads) {
  if (cache_segments < 0) return nullptr;
  auto* e = new mordred_engine();
  e->device = device;
  e->cache_segments = cache_segments;
  if (device < 0) {  // host-only engine: CPU operators, no HIP calls
    e->cache_segments = 0;
    e->pool = std::make_unique<Pool>(threads);
    return e;
  }
  bool ok = hipSetDevice(device) == hipSuccess;
  ok = ok && (cache_segments == 0 ||
              hipMalloc(&e->cache, (size_t)cache_segments * kSeg * sizeof(int)) == hipSuccess);
  ok = ok && hipStreamCreate(&e->s_main) == hipSuccess && hipStreamCreate(&e->s_copy[0]) == hipSuccess &&
       hipStreamCreate(&e->s_copy[1]) == hipSuccess && hipStreamCreate(&e->s_order) == hipSuccess &&
       hipEventCreateWithFlags(&e->ev_built, hipEventDisableTiming) == hipSuccess;
  if (ok) {
    std::vector<short> ident(kMaxIdentity);
    for (int i = 0; i < kMaxIdentity; i++) ident[i] = (short)i;
    ok = hipMalloc(&e->d_identity, kMaxIdentity * sizeof(short)) == hipSuccess &&
         hipMemcpy(e->d_identity, ident.data(), kMaxIdentity * sizeof(short), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok) {
    mordred_engine_destroy(e);
    return nullptr;
  }
  for (int s = cache_segments - 1; s >= 0; s--) e->free_slots.push_back(s);
  e->pool = std::make_unique<Pool>(threads);
  return e;
}

void mordred_engine_destroy(mordred_engine* e) {
  if (!e) return;
  if (e->device < 0) {
    for (auto& kv : e->cols) free_column(e, kv.second.get());
    delete e;
    return;
  }
  (void)hipSetDevice(e->device);
  for (auto& h : e->d_ht)
    if (h.p) (void)mordred_probe_table_unregister(static_cast<const int*>(h.p));
  for (auto& kv : e->cols) free_column(e, kv.second.get());
  e->cols.clear();
  if (e->cm) mordred_cost_model_destroy(e->cm);
  if (e->cm_place) mordred_cost_model_destroy(e->cm_place);
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->ev_built) (void)hipEventDestroy(e->ev_built);
  for (hipStream_t s : {e->s_main, e->s_copy[0], e->s_copy[1], e->s_order})
    if (s) (void)hipStreamDestroy(s);
  if (e->cache) (void)hipFree(e->cache);
  if (e->d_identity) (void)hipFree(e->d_identity);
  delete e;
}

const char* mordred_engine_last_error(const mordred_engine* e) { return e ? e->err.c_str() : "no engine"; }

int mordred_engine_add_column(mordred_engine* e, const char* name, const int* values, long long n,
                              int gpu_resident) {
  if (!e || !name || (!values && n > 0) || n < 0) return e ? fail(e, "bad column") : -1;
  return add_column(e, name, values, -1, n, gpu_resident);
}

int mordred_engine_column_segments(const mordred_engine* e, const char* name) {
  if (!e || !name) return -1;
  auto it = e->cols.find(name);
  return it == e->cols.end() ? -1 : it->second->nseg;
}

int mordred_engine_cache_segment(mordred_engine* e, const char* name, int segment) {
  if (!e) return -1;
  if (e->device < 0) return fail(e, "host-only engine has no segment cache");
  Column* c = col(e, name);
  if (!c) return fail(e, "unknown column");
  HIP_OK(hipSetDevice(e->device));
  return cache_seg(e, c, segment);
}

int mordred_engine_evict_segment(mordred_engine* e, const char* name, int segment) {
  if (!e) return -1;
  if (e->device < 0) return fail(e, "host-only engine has no segment cache");
  Column* c = col(e, name);
  if (!c) return fail(e, "unknown column");
  HIP_OK(hipSetDevice(e->device));
  return evict_seg(e, c, segment);
}

int mordred_engine_is_cached(const mordred_engine* e, const char* name, int segment) {
  if (!e || !name) return -1;
  auto it = e->cols.find(name);
  if (it == e->cols.end() || segment < 0 || segment >= it->second->nseg) return -1;
  return it->second->slot[segment] >= 0;
}

int mordred_engine_set_bandwidth(mordred_engine* e, double bw_cpu, double bw_pci) {
  if (!e || !(bw_cpu > 0) || !(bw_pci > 0)) return e ? fail(e, "bandwidths must be > 0") : -1;
  e->bw_cpu = bw_cpu;
  e->bw_pci = bw_pci;
  return 0;
}

int mordred_engine_load_column_file(mordred_engine* e, const char* name, const char* path, long long n,
                                    int gpu_resident) {
  if (!e) return -1;
  if (!name || !path || n < -1) return fail(e, "bad column file arguments");
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(e, std::string("cannot open column file ") + path);
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    return fail(e, std::string("cannot stat column file ") + path);
  }
  if (n == -1) {
    if (sb.st_size % (off_t)sizeof(int) != 0) {
      close(fd);
      return fail(e, std::string("column file size is not a multiple of 4 bytes: ") + path);
    }
    n = (long long)(sb.st_size / (off_t)sizeof(int));
  } else if ((long long)sb.st_size < n * (long long)sizeof(int)) {
    close(fd);
    return fail(e, std::string("column file shorter than n values: ") + path);
  }
  const int rc = add_column(e, name, nullptr, fd, n, gpu_resident);
  close(fd);
  return rc;
}

int mordred_engine_store_column_file(mordred_engine* e, const char* name, const char* path) {
  if (!e) return -1;
  Column* c = col(e, name);
  if (!c || !path) return fail(e, "unknown column or NULL path");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(e, std::string("cannot create column file ") + path);
  const size_t w = c->n > 0 ? std::fwrite(c->host, sizeof(int), (size_t)c->n, f) : 0;
  const bool ok = w == (size_t)c->n && std::fclose(f) == 0;
  if (!ok) return fail(e, std::string("short write to column file ") + path);
  return 0;
}

int mordred_engine_load_segment_minmax(mordred_engine* e, const char* name, const char* path) {
  if (!e) return -1;
  Column* c = col(e, name);
  if (!c || !path) return fail(e, "unknown column or NULL path");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(e, std::string("cannot open minmax file ") + path);
  std::vector<int> mn, mx;
  char line[128];
  std::string bad;
  while (bad.empty() && std::fgets(line, sizeof line, f)) {
    if (line[0] == '\n' || line[0] == '\0') continue;
    char* p = line;
    char* q = nullptr;
    const long a = std::strtol(p, &q, 10);
    if (q == p) bad = "expected '<min> <max>'";
    p = q;
    const long b = bad.empty() ? std::strtol(p, &q, 10) : 0;
    if (bad.empty() && q == p) bad = "expected '<min> <max>'";
    if (bad.empty() && (a < INT_MIN || a > INT_MAX || b < INT_MIN || b > INT_MAX)) bad = "value out of int32 range";
    if (bad.empty() && a > b) bad = "min > max";
    if (bad.empty()) {
      mn.push_back((int)a);
      mx.push_back((int)b);
    }
  }
  std::fclose(f);
  if (!bad.empty())
    return fail(e, std::string("minmax file ") + path + " line " + std::to_string(mn.size() + 1) + ": " + bad);
  if ((int)mn.size() != c->nseg)
    return fail(e, std::string("minmax file ") + path + " has " + std::to_string(mn.size()) + " lines for " +
                       std::to_string(c->nseg) + " segments");
  // skipping with ranges tighter than the data would drop rows: refuse them
  for (int seg = 0; seg < c->nseg; seg++)
    if (mn[seg] > c->emin[seg] || mx[seg] < c->emax[seg])
      return fail(e, std::string("minmax file ") + path + " segment " + std::to_string(seg) +
                         " does not cover the column's values");
  c->smin = mn;
  c->smax = mx;
  c->clustered = -1;
  return 0;
}

int mordred_engine_store_segment_minmax(mordred_engine* e, const char* name, const char* path) {
  if (!e) return -1;
  Column* c = col(e, name);
  if (!c || !path) return fail(e, "unknown column or NULL path");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(e, std::string("cannot create minmax file ") + path);
  bool ok = true;
  for (int seg = 0; seg < c->nseg && ok; seg++) ok = std::fprintf(f, "%d %d\n", c->smin[seg], c->smax[seg]) > 0;
  ok = std::fclose(f) == 0 && ok;
  if (!ok) return fail(e, std::string("short write to minmax file ") + path);
  return 0;
}

int mordred_engine_segment_minmax(mordred_engine* e, const char* name, int* mins, int* maxs) {
  if (!e) return -1;
  Column* c = col(e, name);
  if (!c || !mins || !maxs) return fail(e, "unknown column or NULL output");
  std::copy(c->smin.begin(), c->smin.end(), mins);
  std::copy(c->smax.begin(), c->smax.end(), maxs);
  return 0;
}

int mordred_engine_set_skipping(mordred_engine* e, int on) {
  if (!e) return -1;
  e->skipping = on != 0;
  return 0;
}

int mordred_engine_set_batch_sharing(mordred_engine* e, int on) {
  if (!e) return -1;
  e->batch_share = on != 0;
  return 0;
}

int mordred_engine_set_batch_forms(mordred_engine* e, int forms) {
  if (!e) return -1;
  if (forms & ~MORDRED_BATCH_FORM_ALL) return fail(e, "unknown batch form bits");
  e->batch_forms = forms;
  return 0;
}

long long mordred_engine_struct_size(const char* name) {
  if (!name) return 0;
  const std::string n(name);
#define MORDRED_SIZE_OF(T) \
  if (n == #T) return (long long)sizeof(T);
  MORDRED_SIZE_OF(mordred_join)
  MORDRED_SIZE_OF(mordred_fact_filter)
  MORDRED_SIZE_OF(mordred_star_query)
  MORDRED_SIZE_OF(mordred_run_stats)
  MORDRED_SIZE_OF(mordred_placement_stats)
  MORDRED_SIZE_OF(mordred_batch_config)
  MORDRED_SIZE_OF(mordred_batch_stats)
  MORDRED_SIZE_OF(mordred_batch_tune_entry)
  MORDRED_SIZE_OF(mordred_serve_opts)
  MORDRED_SIZE_OF(mordred_serve_times)
  MORDRED_SIZE_OF(mordred_serve_stats)
#undef MORDRED_SIZE_OF
  return 0;
}

int mordred_engine_set_operator_order(mordred_engine* e, int order) {
  if (!e) return -1;
  if (order < MORDRED_ORDER_GPU_FIRST || order > MORDRED_ORDER_AUTO) return fail(e, "bad operator order");
  e->op_order = order;
  return 0;
}

int mordred_engine_set_prefix_streaming(mordred_engine* e, int on) {
  if (!e) return -1;
  e->prefix_stream = on != 0;
  return 0;
}

int mordred_engine_place(mordred_engine* e, const mordred_star_query* q, double fraction) {
  if (!e) return -1;
  if (e->device < 0) return fail(e, "host-only engine has no segment cache");
  Resolved r;
  int rc = resolve(e, q, r);
  if (rc) return rc;
  HIP_OK(hipSetDevice(e->device));
  // permute_cost weight of a segment: what moving its pipeline from the CPU to
  // the GPU saves (CostModel.cu:147-181), accumulated over calls like the
  // reference's segment weights (CacheManager.cu:533-587).
  const Selectivity sel = estimate_selectivity(r);
  // segments the query skips weigh nothing and are ranked last (the
  // reference updates segment statistics only for segments it processes,
  // QueryOptimizer.cu:1696-1699)
  std::vector<char> skip;
  if (e->skipping) skip_segments(e, r, skip);
  // With the learned plan model ready, a segment weighs what the model
  // predicts caching one more segment saves this query (its COST_SPLIT run
  // with the analytic split, before and after) instead of the analytic CPU
  // cost; the weights accumulate over calls either way.
  double learned_saving = -1;  // ms per full segment
  if (place_model_ready(e)) {
    int cached = 0, n = 0;
    for (int s = 0; s < r.nseg; s++) {
      if (!skip.empty() && skip[s]) continue;
      bool all = true;
      for (Column* c : r.fact_cols) all &= c->slot[s] >= 0;
      (all ? cached : n)++;
    }
    if (n > 0) {
      const double t_cpu = cpu_cost(r, sel, kSeg, e->bw_cpu) * 1e3;
      const double t_str = transfer_cost(r, kSeg, e->bw_pci) * 1e3;
      long long dim_rows = 0;
      for (int k = 0; k < 4; k++)
        if (r.join[k]) dim_rows += r.dkey[k]->n;
      const double t_fix = e->build_ns_per_row * (double)dim_rows * 1e-6;
      float xs[2 * MORDRED_COST_FEATURES], pred[2];
      for (int c = 0; c < 2; c++) {
        const int nn = n - c;
        split_features(e, r, sel, cached + c, nn, split_balance(nn, t_cpu, t_str, 0, t_fix), t_cpu, t_str, 0, t_fix,
                       -1, 0.0, 0.0, xs + c * MORDRED_COST_FEATURES);
      }
      if (mordred_cost_model_predict(e->cm_place, xs, 2, pred) == 0) {
        learned_saving = std::max(0.0, std::exp((double)pred[0]) - std::exp((double)pred[1]));
        e->learned_residency_calls++;
      }
    }
  }
  std::vector<std::pair<double, int>> rank;
  for (int s = 0; s < r.nseg; s++) {
    if (!skip.empty() && skip[s]) {
      rank.emplace_back(1.0, s);
      continue;
    }
    const double unit = learned_saving >= 0 ? learned_saving * (double)segment_rows(r, s) / kSeg
                                            : cpu_cost(r, sel, (double)segment_rows(r, s), e->bw_cpu);
    double w = unit;
    for (Column* c : r.fact_cols) {
      const long long id = ((long long)(intptr_t)c << 16) ^ s;
      w += (e->seg_weight[id] += unit);
    }
    rank.emplace_back(-w, s);
  }
  std::sort(rank.begin(), rank.end());
  const int want = (int)std::min<double>(r.nseg, std::max(0.0, fraction) * r.nseg + 0.5);
  std::vector<char> keep(r.nseg, 0);
  for (int i = 0; i < want; i++) keep[rank[i].second] = 1;
  for (int s = 0; s < r.nseg; s++)
    if (!keep[s])
      for (Column* c : r.fact_cols)
        if ((rc = evict_seg(e, c, s))) return rc;
  for (int s = 0; s < r.nseg; s++)
    if (keep[s])
      for (Column* c : r.fact_cols)
        if ((rc = cache_seg(e, c, s))) return rc;
  return 0;
}

int mordred_engine_run(mordred_engine* e, const mordred_star_query* q, int placement, int* res,
                       mordred_run_stats* stats) {
  if (!e) return -1;
  if (!res) return fail(e, "res is NULL");
  if (placement < MORDRED_PLACE_RESIDENCY || placement > MORDRED_PLACE_CPU_ONLY) return fail(e, "bad placement");
  if (e->device < 0 && placement != MORDRED_PLACE_CPU_ONLY) return fail(e, "host-only engine runs CPU_ONLY");
  Resolved r;
  int rc = resolve(e, q, r);
  if (rc) return rc;
  for (int k = 0; k < 4; k++) {
    if (!r.join[k]) continue;
    for (Column* c : {r.dkey[k], r.dfilt[k], r.dval[k]})
      if (c && !c->pinned_gpu && placement != MORDRED_PLACE_CPU_ONLY)
        return fail(e, "dimension columns must be GPU-resident (add_column gpu_resident=1)");
  }
  const auto t0 = Clock::now();
  const int G = q->kind == MORDRED_KIND_GROUPBY ? q->total_val : 1;
  const size_t res_ints = 6 * (size_t)G;
  std::memset(res, 0, res_ints * sizeof(int));
  mordred_run_stats st{};

  // ---- segment groups by residency (QueryOptimizer.cu:1680-1716), without
  // the segments the predicates exclude
  std::vector<char> skip;
  if (e->skipping) st.skipped_segments = skip_segments(e, r, skip);
  GpuPart g;
  std::vector<int> other;
  for (int s = 0; s < r.nseg; s++) {
    if (!skip.empty() && skip[s]) continue;
    if (placement != MORDRED_PLACE_CPU_ONLY && segment_cached(r, s)) g.gpu_segs.push_back(s);
    else other.push_back(s);
  }
  // partly cached segments, grouped by which of the query's filters and
  // probes have their fact column cached (candidates for a mixed pipeline)
  std::vector<MixedGroup> mixed;
  if (placement == MORDRED_PLACE_RESIDENCY || placement == MORDRED_PLACE_COST_SPLIT) {
    std::vector<int> plain;
    for (int s : other) {
      MixedGroup pat;
      bool any = false;
      for (int i = 0; i < r.q->nfilters; i++) any |= pat.gpu_filter[i] = r.filt[i]->slot[s] >= 0;
      for (int k = 0; k < 4; k++) any |= pat.gpu_probe[k] = r.join[k] && r.key[k]->slot[s] >= 0;
      if (!any) {
        plain.push_back(s);
        continue;
      }
      MixedGroup* grp = nullptr;
      for (MixedGroup& m : mixed)
        if (std::equal(m.gpu_filter, m.gpu_filter + 2, pat.gpu_filter) &&
            std::equal(m.gpu_probe, m.gpu_probe + 4, pat.gpu_probe))
          grp = &m;
      if (!grp) {
        mixed.push_back(pat);
        grp = &mixed.back();
      }
      grp->segs.push_back(s);
    }
    other.swap(plain);
    for (MixedGroup& m : mixed) m.rows = group_rows(r, m.segs);
  }
  const bool may_stream = !other.empty() &&
                          (placement == MORDRED_PLACE_GPU_ONLY || placement == MORDRED_PLACE_COST_SPLIT);
  const bool use_gpu = !g.gpu_segs.empty() || may_stream || !mixed.empty();
  const size_t staging = 2 * (size_t)kStreamBatch * r.fact_cols.size();

  // ---- GPU half on its own thread (builds + cached group start now)
  std::promise<SplitPlan> streamed;
  g.streamed = streamed.get_future();
  MixedFeed feed;
  g.feed = &feed;
  std::future<int> built = g.built.get_future();
  std::thread feeder;
  if (use_gpu) feeder = std::thread([&] { g.rc = gpu_part(e, r, g, res_ints); });

  // ---- selectivities (sampled, so the plan does not wait for the CPU
  // hash tables, which are built once the GPU side has its plan)
  const auto tc0 = Clock::now();
  CpuTables t;
  const Selectivity sel = estimate_selectivity(r);
  set_probe_order(r, sel, t);
  bool cpu_tables = false;

  // ---- split the non-cached segments between the CPU and the PCIe stream
  std::vector<int> cpu_segs, stream_segs;
  const double t_cpu = cpu_cost(r, sel, kSeg, e->bw_cpu) * 1e3;
  // the CPU hash-table builds, at the engine's measured time per dimension row
  long long dim_rows = 0;
  for (int k = 0; k < 4; k++)
    if (r.join[k]) dim_rows += r.dkey[k]->n;
  const double t_fix = e->build_ns_per_row * (double)dim_rows * 1e-6;
  st.est_cpu_build_ms = t_fix;
  const double t_str = transfer_cost(r, kSeg, e->bw_pci) * 1e3;
  st.est_cpu_ms_per_segment = t_cpu;
  st.est_stream_ms_per_segment = t_str;
  if (placement == MORDRED_PLACE_COST_SPLIT && !mixed.empty()) {
    // a mixed group stays mixed when the analytic model times it below both
    // whole-segment plans: the GPU operators are free (CostModel.cu:283-363),
    // the survivors' offsets cross PCIe, and the CPU operators touch one cache
    // line per survivor and column (random access)
    std::vector<MixedGroup> keep;
    for (MixedGroup& m : mixed) {
      const double t_mix = mixed_cost(e, r, sel, m, false).total();
      st.est_mixed_ms_per_segment = t_mix;
      if (t_mix < std::min(t_cpu, t_str)) keep.push_back(std::move(m));
      else other.insert(other.end(), m.segs.begin(), m.segs.end());
    }
    mixed.swap(keep);
    std::sort(other.begin(), other.end());
  }
  // prefix streaming: of a whole non-resident segment, only the columns of a
  // subset of the fact filters and probes cross PCIe; those operators run on
  // the GPU and the CPU operators finish the survivors, as in a mixed
  // pipeline (the survivors' CPU operators run on the CPU side of the split)
  MixedGroup pgroup;
  std::vector<int> pstream_segs;
  bool prefix_plan = false;
  if (placement == MORDRED_PLACE_COST_SPLIT && e->prefix_stream && !other.empty()) {
    // the GPU operator set: every non-empty subset of the fact filters and
    // probes (at most 6 operators), timed by the split model; the best one
    // is kept if it beats whole-segment streaming
    int ops[6], nops = 0;
    for (int i = 0; i < r.q->nfilters; i++) ops[nops++] = i;
    for (int k = 0; k < 4; k++)
      if (r.join[k]) ops[nops++] = 2 + k;
    const int n = (int)other.size();
    const int kw = split_balance(n, t_cpu, t_str, 0, t_fix);
    double best = split_analytic_ms(n, kw, t_cpu, t_str, 0, t_fix);
    for (int set = 1; set < (1 << nops); set++) {
      MixedGroup cand;
      for (int b = 0; b < nops; b++) {
        if (!(set >> b & 1)) continue;
        if (ops[b] < 2) cand.gpu_filter[ops[b]] = true;
        else cand.gpu_probe[ops[b] - 2] = true;
      }
      const MixedCost pc = mixed_cost(e, r, sel, cand, true);
      const int kp = split_balance(n, t_cpu, pc.gpu, pc.cpu, t_fix);
      const double tp = split_analytic_ms(n, kp, t_cpu, pc.gpu, pc.cpu, t_fix);
      if (tp < best) {
        best = tp;
        pgroup = cand;
        prefix_plan = true;
        st.est_prefix_ms_per_segment = pc.gpu;
        st.est_prefix_cpu_ms_per_segment = pc.cpu;
      }
    }
  }
  // the GPU side's time per segment, and the CPU time it leaves per segment
  const double t_gpu = prefix_plan ? st.est_prefix_ms_per_segment : t_str;
  const double t_gcpu = prefix_plan ? st.est_prefix_cpu_ms_per_segment : 0.0;
  // the operator order of the mixed groups (the rule: the engine's order, or
  // under MORDRED_ORDER_AUTO the fastest measured so far for this query and
  // residency); another order than GPU-first runs on this thread (below)
  std::string order_key;
  int order = e->op_order;
  if (order == MORDRED_ORDER_AUTO) {
    order = MORDRED_ORDER_GPU_FIRST;
    if (!mixed.empty()) {
      order_key = order_signature(r, mixed);
      if (e->order_ms.size() >= 4096 && !e->order_ms.count(order_key)) e->order_ms.clear();  // bounded memory
      const std::array<double, 3>& ms = e->order_ms[order_key];
      int best = -1;
      for (int o = 0; o < 3 && best < 0; o++)
        if (ms[o] == 0) best = o;  // not tried yet
      if (best < 0) best = (int)(std::min_element(ms.begin(), ms.end()) - ms.begin());
      order = best;
    }
  }
  long long mixed_segs = 0;
  double mixed_ms = 0;  // the analytic model's time of the mixed groups
  for (const MixedGroup& m : mixed) {
    mixed_segs += (long long)m.segs.size();
    mixed_ms += mixed_cost(e, r, sel, m, false).total() * (double)m.segs.size();
  }
  const double mixed_frac = r.nseg > 0 ? (double)mixed_segs / r.nseg : 0.0;
  const int n = (int)other.size();
  int k = placement == MORDRED_PLACE_GPU_ONLY ? n
          : placement == MORDRED_PLACE_COST_SPLIT ? split_balance(n, t_cpu, t_gpu, t_gcpu, t_fix) : 0;
  // the learned plan model: candidate plans are the split k (COST_SPLIT:
  // k = n*c/8 for c = 0..8, and the analytic k) times the operator order
  // (all three when the engine's order is AUTO and mixed groups run, else
  // the rule's); the model's fastest plan runs while it is trusted
  const bool split_plan = placement == MORDRED_PLACE_COST_SPLIT && n > 0;
  const bool order_plan = !mixed.empty() && e->op_order == MORDRED_ORDER_AUTO;
  const bool model_plans = e->cm_place && e->place_mode > 0 && (split_plan || order_plan);
  if (model_plans) {
    std::vector<int> ks{k}, os{order};
    if (split_plan)
      for (int c = 0; c <= 8; c++) ks.push_back((int)((long long)n * c / 8));
    if (order_plan) os = {MORDRED_ORDER_GPU_FIRST, MORDRED_ORDER_CPU_FIRST, MORDRED_ORDER_ALTERNATE};
    std::sort(ks.begin(), ks.end());
    ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
    std::vector<std::pair<int, int>> plans;
    for (int kk : ks)
      for (int o : os) plans.emplace_back(kk, o);
    std::vector<float> xs(plans.size() * MORDRED_COST_FEATURES), pred(plans.size());
    for (size_t c = 0; c < plans.size(); c++)
      split_features(e, r, sel, (int)g.gpu_segs.size(), n, plans[c].first, t_cpu, t_gpu, t_gcpu, t_fix,
                     plans[c].second, mixed_frac, mixed_ms, &xs[c * MORDRED_COST_FEATURES]);
    // a model error leaves the rule's plan (the GPU half already runs)
    if (place_model_ready(e) && mordred_cost_model_predict(e->cm_place, xs.data(), (int)plans.size(), pred.data()) == 0) {
      size_t best = 0;
      for (size_t c = 1; c < plans.size(); c++)
        if (pred[c] < pred[best]) best = c;
      k = plans[best].first;
      if (order_plan && plans[best].second != order) e->learned_order_choices++;
      order = plans[best].second;
      st.plan_model = 1;
      st.plan_predicted_ms = std::exp((double)pred[best]);
    }
    if (e->place_mode == MORDRED_LEARNED_EXPLORE) {
      // exploration: one plan in four is drawn at random
      e->place_rng ^= e->place_rng << 13;
      e->place_rng ^= e->place_rng >> 7;
      e->place_rng ^= e->place_rng << 17;
      if ((e->place_rng & 3) == 0) {
        const auto& pl = plans[(size_t)((e->place_rng >> 8) % plans.size())];
        k = pl.first;
        order = pl.second;
        st.plan_model = 2;
      }
    }
  }
  st.plan_analytic_ms = split_analytic_ms(n, k, t_cpu, t_gpu, t_gcpu, t_fix) + mixed_ms;
  st.plan_streamed = k;
  if (placement == MORDRED_PLACE_GPU_ONLY) {
    stream_segs = other;
  } else if (placement == MORDRED_PLACE_COST_SPLIT) {
    // both sides run concurrently: k segments go to the GPU side (so that
    // both finish together, by the analytic model)
    if (prefix_plan) pstream_segs.assign(other.begin(), other.begin() + k);
    else stream_segs.assign(other.begin(), other.begin() + k);
    cpu_segs.assign(other.begin() + k, other.end());
  } else {
    cpu_segs = other;
  }
  if ((!stream_segs.empty() || !pstream_segs.empty()) && e->free_slots.size() < staging) {
    // not enough free cache slots for two staging batches: CPU instead
    cpu_segs.insert(cpu_segs.end(), stream_segs.begin(), stream_segs.end());
    cpu_segs.insert(cpu_segs.end(), pstream_segs.begin(), pstream_segs.end());
    stream_segs.clear();
    pstream_segs.clear();
  }
  std::sort(cpu_segs.begin(), cpu_segs.end());
  const bool ordered = order != MORDRED_ORDER_GPU_FIRST && !mixed.empty();
  std::vector<MixedGroup> no_mixed;
  // the prefix-streamed columns cross PCIe packed: their segments are packed
  // once, here, before the GPU side (which waits for this plan) streams them
  if (!pstream_segs.empty()) {
    for (int i = 0; i < r.q->nfilters; i++)
      if (pgroup.gpu_filter[i]) pack_segments(e, r.filt[i], pstream_segs);
    for (int k = 0; k < 4; k++)
      if (r.join[k] && pgroup.gpu_probe[k]) pack_segments(e, r.key[k], pstream_segs);
  }
  streamed.set_value(SplitPlan{stream_segs, ordered ? &no_mixed : &mixed, pstream_segs, &pgroup});

  // ---- CPU hash tables, for every CPU operator of the run
  if (!cpu_segs.empty() || !mixed.empty() || !pstream_segs.empty()) {
    const auto tb0 = Clock::now();
    cpu_build(e, r, t);
    cpu_tables = true;
    st.cpu_build_ms = ms_since(tb0);
    if (dim_rows >= (1 << 16)) {
      const double ns = st.cpu_build_ms * 1e6 / (double)dim_rows;
      e->build_ns_per_row = e->build_ns_per_row > 0 ? 0.7 * e->build_ns_per_row + 0.3 * ns : ns;
    }
  }

  // ---- CPU: the CPU segment group, concurrently with the GPU
  std::vector<int> cpu_res;
  if (!cpu_segs.empty()) {
    cpu_res.assign(res_ints, 0);
    cpu_fact(e, r, t, cpu_segs, cpu_res.data());
    st.cpu_rows = group_rows(r, cpu_segs);
    st.cpu_segments = (int)cpu_segs.size();
  }
  // ---- mixed groups: the CPU operators over the GPU operators' survivors
  std::vector<int> ord_res;  // GPU group-by stages of ordered mixed groups
  st.operator_order = order;
  if (!mixed.empty() || !pstream_segs.empty()) {
    if (cpu_res.empty()) cpu_res.assign(res_ints, 0);
    std::vector<CpuAcc> acc(e->pool->size());
    if (ordered) {
      int rc = built.get() == 0 ? 0 : -1;
      if (!rc) {
        const hipError_t he = e->d_res_ord.reserve(res_ints * sizeof(int));
        rc = he != hipSuccess ? hip_fail(e, he, "ordered pipelines: result buffer")
                              : order_groups(e, r, t, mixed, order, static_cast<int*>(e->d_res_ord.p), res_ints, acc,
                                             st);
      }
      if (!rc) {
        ord_res.assign(res_ints, 0);
        const hipError_t he = hipMemcpy(ord_res.data(), e->d_res_ord.p, res_ints * sizeof(int), hipMemcpyDeviceToHost);
        if (he != hipSuccess) rc = hip_fail(e, he, "ordered pipelines: result copy");
      }
      if (rc) {
        feed.finish(false);
        if (feeder.joinable()) feeder.join();
        return g.rc ? g.rc : rc;
      }
    }
    MixedChunk ch;
    while (feed.pop(ch)) {
      const auto ts0 = Clock::now();
      cpu_suffix(e, r, t, ch.group < 0 ? pgroup : mixed[ch.group], ch, acc);
      st.suffix_ms += ms_since(ts0);
    }
    cpu_combine(r, acc, cpu_res.data());
    if (!feed.ok) {
      if (feeder.joinable()) feeder.join();
      return g.rc ? g.rc : fail(e, "mixed pipeline: GPU part failed");
    }
    double lines = (double)pgroup.survivors * mixed_lines(r, pgroup);
    for (const MixedGroup& m : mixed) {
      st.mixed_segments += (int)m.segs.size();
      st.mixed_rows += m.rows;
      st.mixed_survivors += m.survivors;
      if (!ordered) lines += (double)m.survivors * mixed_lines(r, m);
    }
    // calibrate the CPU operators' time per survivor line (moving average)
    if (lines >= (1 << 18) && st.suffix_ms > 0) {
      const double ns = st.suffix_ms * 1e6 / lines;
      e->suffix_ns_per_line = e->suffix_ns_per_line > 0 ? 0.7 * e->suffix_ns_per_line + 0.3 * ns : ns;
    }
  }
  st.cpu_ms = cpu_tables ? ms_since(tc0) : 0.0;

  // ---- wait for the GPU, merge
  if (feeder.joinable()) feeder.join();
  if (g.rc) return g.rc;
  st.gpu_ms = g.ms;
  st.gpu_rows = g.gpu_rows;
  st.gpu_segments = g.gpu_segments;
  st.stream_rows = g.stream_rows;
  st.stream_segments = g.stream_segments;
  st.prefix_segments = g.prefix_segments;
  st.prefix_survivors = pgroup.survivors;
  st.prefix_packed_bytes = g.packed_bytes;
  st.prefix_rows = g.prefix_rows;
  if (!ord_res.empty()) {
    if (g.res.empty()) g.res.swap(ord_res);
    else if (q->kind == MORDRED_KIND_GROUPBY) merge(g.res.data(), ord_res.data(), G);
    else reinterpret_cast<long long*>(g.res.data())[2] += reinterpret_cast<long long*>(ord_res.data())[2];
  }
  const auto tm0 = Clock::now();
  if (!cpu_res.empty() && !g.res.empty()) {
    if (q->kind == MORDRED_KIND_GROUPBY) merge(cpu_res.data(), g.res.data(), G);
    else reinterpret_cast<long long*>(cpu_res.data())[2] += reinterpret_cast<long long*>(g.res.data())[2];
    std::memcpy(res, cpu_res.data(), res_ints * sizeof(int));
  } else if (!cpu_res.empty()) {
    std::memcpy(res, cpu_res.data(), res_ints * sizeof(int));
  } else if (!g.res.empty()) {
    std::memcpy(res, g.res.data(), res_ints * sizeof(int));
  }
  st.merge_ms = ms_since(tm0);
  st.total_ms = ms_since(t0);
  // learn from the measured run (a plan the model was given, as it ran)
  if (model_plans && (!split_plan || st.plan_streamed == st.stream_segments + st.prefix_segments)) {
    float f[MORDRED_COST_FEATURES];
    split_features(e, r, sel, st.gpu_segments, n, split_plan ? st.plan_streamed : k, t_cpu, t_gpu, t_gcpu, t_fix,
                   order, mixed_frac, mixed_ms, f);
    if (mordred_cost_model_observe(e->cm_place, f, std::max(1e-3, st.total_ms), std::max(1e-3, st.plan_analytic_ms)))
      return fail(e, std::string("placement model: ") + mordred_cost_model_last_error(e->cm_place));
    if (++e->place_since_train >= 8) {
      e->place_since_train = 0;
      if (mordred_cost_model_train(e->cm_place, 32, 64, 2e-3f, nullptr))
        return fail(e, std::string("placement model: ") + mordred_cost_model_last_error(e->cm_place));
      e->place_trains++;
    }
  }
  if (!order_key.empty()) {
    double& slot = e->order_ms[order_key][order];
    slot = slot > 0 ? std::min(slot, st.total_ms) : std::max(st.total_ms, 1e-6);
  }
  if (stats) *stats = st;
  return 0;
}

int mordred_engine_set_learned_placement(mordred_engine* e, int mode) {
  if (!e) return -1;
  if (mode < MORDRED_LEARNED_OFF || mode > MORDRED_LEARNED_FORCE) return fail(e, "bad learned placement mode");
  if (mode != MORDRED_LEARNED_OFF && e->device < 0) return fail(e, "the learned placement model needs a GPU");
  if (mode != MORDRED_LEARNED_OFF && !e->cm_place) {
    e->cm_place = mordred_cost_model_create(e->device, 777);
    if (!e->cm_place) return fail(e, "placement model creation failed");
  }
  e->place_mode = mode;
  return 0;
}

int mordred_engine_placement_model_stats(mordred_engine* e, mordred_placement_stats* st) {
  if (!e || !st) return e ? fail(e, "stats is NULL") : -1;
  *st = mordred_placement_stats{};
  if (!e->cm_place) return 0;
  mordred_cost_model_stats cs{};
  const int rc = mordred_cost_model_get_stats(e->cm_place, &cs);
  if (rc) return fail(e, "placement model stats failed");
  st->samples = cs.samples;
  st->scored = cs.scored;
  st->learned_error = cs.ewma_abs_log_error;
  st->analytic_error = cs.ewma_abs_log_error_analytic;
  st->trusted = cs.trusted;
  st->order_choices = e->learned_order_choices;
  st->residency_calls = e->learned_residency_calls;
  return 0;
}

int mordred_engine_batch(mordred_engine* e, const mordred_star_query* qs, int nq, int* const* res,
                         const mordred_batch_config* cfg, mordred_batch_stats* stats) {
  if (!e) return -1;
  if (e->device < 0) return fail(e, "host-only engine has no batch path");
  if (!qs || !res) return fail(e, "queries / results are NULL");
  for (int i = 0; i < nq; i++)
    if (!res[i]) return fail(e, "a result pointer is NULL");
  const mordred::BatchConfig c = cfg ? mordred::BatchConfig{cfg->block, cfg->vpt, cfg->lds_bytes} : e->batch_cfg;
  HIP_OK(hipSetDevice(e->device));
  std::vector<const mordred_star_query*> qp(std::max(0, nq));
  for (int i = 0; i < nq; i++) qp[i] = &qs[i];
  BatchPlan b;
  const int rc = plan_batch(e, qp.data(), nq, c, b);
  if (rc) return rc;
  return run_batch(e, b, c, res, stats);
}

int mordred_engine_batch_tune(mordred_engine* e, const mordred_star_query* qs, int nq, int reps,
                              mordred_batch_config* best, mordred_batch_tune_entry* table, int max_entries,
                              int* nentries) {
  if (!e) return -1;
  if (e->device < 0) return fail(e, "host-only engine has no batch path");
  if (!qs || nq < 1 || nq > mordred::kBatchMaxQ) return fail(e, "a batch holds 1..64 queries");
  HIP_OK(hipSetDevice(e->device));
  reps = std::max(1, reps);
  // candidates: threads per workgroup x items per thread x LDS budget (which
  // decides how many group arrays live in LDS rather than HBM)
  std::vector<mordred::BatchConfig> cands;
  for (int block : {256, 512, 1024})
    for (int vpt : {1, 2})
      for (int lds : {16 << 10, 48 << 10, 96 << 10, 160 << 10}) {
        const mordred::BatchConfig c{block, vpt, lds};
        if (mordred::batch_config_ok(c)) cands.push_back(c);
      }
  std::vector<std::vector<int>> ref, cur;
  std::vector<int*> ptrs(nq);
  double best_ms = -1;
  mordred::BatchConfig best_c = e->batch_cfg;
  int n_out = 0;
  std::string first_err;
  std::vector<const mordred_star_query*> qp(nq);
  for (int i = 0; i < nq; i++) qp[i] = &qs[i];
  for (const auto& c : cands) {
    BatchPlan b;
    double ms = -1;
    int rc = plan_batch(e, qp.data(), nq, c, b);
    if (!rc) {
      cur.assign(nq, {});
      for (int i = 0; i < nq; i++) {
        cur[i].assign(6 * (size_t)b.G[i], 0);
        ptrs[i] = cur[i].data();
      }
      mordred_batch_stats st{};
      rc = run_batch(e, b, c, ptrs.data(), &st);  // warm-up (and capture)
      for (int r = 0; r < reps && !rc; r++) {
        rc = run_batch(e, b, c, ptrs.data(), &st);
        if (!rc && (ms < 0 || st.gpu_ms < ms)) ms = st.gpu_ms;
      }
    }
    if (rc) {
      if (first_err.empty()) first_err = e->err;
      ms = -1;
      (void)hipGetLastError();
    } else if (ref.empty()) {
      ref = cur;
    } else if (ref != cur) {
      return fail(e, "batch autotuner: launch configurations disagree on a result");
    }
    if (table && n_out < max_entries)
      table[n_out++] = mordred_batch_tune_entry{{c.block, c.vpt, c.lds_bytes}, ms};
    if (ms >= 0 && (best_ms < 0 || ms < best_ms)) {
      best_ms = ms;
      best_c = c;
    }
  }
  if (nentries) *nentries = n_out;
  if (best_ms < 0) return fail(e, "batch autotuner: no configuration ran (" + first_err + ")");
  e->batch_cfg = best_c;
  e->batch_tuned = true;
  if (best) *best = mordred_batch_config{best_c.block, best_c.vpt, best_c.lds_bytes};
  return 0;
}

int mordred_engine_serve(mordred_engine* e, const mordred_star_query* qs, int n, const double* arrival_ms,
                         int* const* res, const mordred_serve_opts* o, mordred_serve_times* times,
                         mordred_serve_stats* stats) {
  using namespace mordred;
  if (!e) return -1;
  if (e->device < 0) return fail(e, "host-only engine has no request loop");
  if (!qs || !arrival_ms || !res || !o || !times || n < 0) return fail(e, "serve: NULL argument");
  for (int i = 0; i < n; i++) {
    if (!res[i]) return fail(e, "serve: a result pointer is NULL");
    if (i && arrival_ms[i] < arrival_ms[i - 1]) return fail(e, "serve: arrivals must be ascending");
  }
  HIP_OK(hipSetDevice(e->device));
  mordred_serve_stats st{};
  auto wall0 = Clock::now();
  auto now_ms = [&] { return ms_since(wall0); };
  auto wait_until = [&](double t) {
    while (now_ms() < t) std::this_thread::sleep_for(std::chrono::microseconds(50));
  };
  for (int i = 0; i < n; i++) times[i] = mordred_serve_times{arrival_ms[i], -1, -1, 0};

  if (o->mode == MORDRED_SERVE_PLAIN) {
    for (int i = 0; i < n; i++) {
      wait_until(arrival_ms[i]);
      times[i].admit_ms = now_ms();
      mordred_run_stats rs{};
      const int rc = mordred_engine_run(e, &qs[i], MORDRED_PLACE_RESIDENCY, res[i], &rs);
      if (rc) return rc;
      if (rs.cpu_rows) return fail(e, "serve: fact columns must be cached in HBM");
      times[i].done_ms = now_ms();
      times[i].launches = 1;
      st.launches++;
      st.gpu_ms += rs.gpu_ms;
      st.max_active_seen = 1;
    }
    st.wall_ms = now_ms();
    if (stats) *stats = st;
    return 0;
  }
  if (o->mode != MORDRED_SERVE_CONTINUOUS) return fail(e, "serve: unknown mode");
  const int nslot = std::max(1, std::min(o->max_active > 0 ? o->max_active : kBatchMaxQ, kBatchMaxQ));
  const BatchConfig cfg = o->config.block > 0 ? BatchConfig{o->config.block, o->config.vpt, o->config.lds_bytes}
                                              : e->batch_cfg;
  if (!batch_config_ok(cfg)) return fail(e, "serve: bad batch configuration");
  // per-request estimates and the slot arena (result rows)
  std::vector<QueryEst> est(n);
  int gmax = 1;
  for (int i = 0; i < n; i++) {
    const int rc = estimate_query(e, qs[i], est[i]);
    if (rc) return rc;
    gmax = std::max(gmax, est[i].G);
  }
  SlotArenas ar;
  ar.res_stride = ((6 * (long long)gmax) + 3) & ~3ll;
  HIP_OK(e->d_serve_res.reserve((size_t)nslot * ar.res_stride * sizeof(int)));
  ar.res = static_cast<int*>(e->d_serve_res.p);
  ar.touch_stride = batch_touch_words(gmax, 0);
  HIP_OK(e->d_serve_touch.reserve(std::max<size_t>(16, (size_t)nslot * ar.touch_stride * sizeof(unsigned))));
  ar.touch = static_cast<unsigned*>(e->d_serve_touch.p);
  auto state_bytes = [&](int i) { return 24.0 * est[i].G; };
  if (o->learned && !e->cm) {
    e->cm = mordred_cost_model_create(e->device, 12345);
    if (!e->cm) return fail(e, "serve: cost model creation failed");
  }
  HIP_OK(hipSetDevice(e->device));
  // the fact table: every request reads the same one
  long long nrows = 0;
  int nseg = 0;
  if (n > 0) {
    Resolved r0;
    int rc = resolve(e, &qs[0], r0);
    if (rc) return rc;
    nrows = r0.n;
    nseg = r0.nseg;
  }
  struct Slot {
    int req = -1;
    int start = 0;     // segment where its scan began
    int covered = 0;   // segments scanned for it
  };
  std::vector<Slot> slots(nslot);
  std::vector<const mordred_star_query*> qp(nslot, nullptr);
  int next = 0, done = 0, cur = 0, active = 0;
  double state = 0;
  bool changed = true;
  const int sizes[] = {8, 16, 32, 64, 128, 256, 512, 1024};
  float feat[MORDRED_COST_FEATURES];
  long long since_train = 0;
  auto rows_of = [&](int from, int to) {
    return std::min<long long>((long long)to << kSegLog, nrows) - ((long long)from << kSegLog);
  };
  // One launch of the shared scan.  The loop plans launch k+1 while launch k
  // runs: which queries launch k completes is known when it is planned (each
  // query sees every segment once, so it finishes when its covered segments
  // reach nseg), their slots are free for the next plan, and their result
  // rows are copied to pinned staging right behind launch k on the stream,
  // before launch k+1's fill kernel can reset a reused slot.
  struct Launch {
    bool on = false;
    BatchPlan b;
    std::vector<QueryEst> es;
    int nreset = 0;
    std::vector<std::pair<int, int>> retire;  // (slot, request) that this launch completes
  };
  // pinned staging: the status word, then one result area per slot
  HIP_OK(e->h_arena.reserve(64 + (size_t)nslot * ar.res_stride * sizeof(int)));
  char* const stage = static_cast<char*>(e->h_arena.p);
  auto stage_of = [&](int sidx) { return stage + 64 + (size_t)sidx * ar.res_stride * sizeof(int); };
  auto result_ints = [&](int req) {
    return 6 * (size_t)(qs[req].kind == MORDRED_KIND_GROUPBY ? qs[req].total_val : 1);
  };
  auto launch = [&](Launch& L) -> int {
    int rc = launch_batch(e, L.b, cfg, nullptr);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(stage, e->d_status.p, sizeof(unsigned), hipMemcpyDeviceToHost, e->s_main));
    for (const auto& [sidx, req] : L.retire)
      HIP_OK(hipMemcpyAsync(stage_of(sidx), ar.res + sidx * ar.res_stride, result_ints(req) * sizeof(int),
                            hipMemcpyDeviceToHost, e->s_main));
    return 0;
  };
  auto complete = [&](Launch& L) -> int {
    HIP_OK(hipStreamSynchronize(e->s_main));
    L.on = false;
    unsigned collisions = 0;
    std::memcpy(&collisions, stage, sizeof collisions);
    int rc = batch_status(e, collisions);
    if (rc) return rc;
    float gms = 0;
    HIP_OK(hipEventElapsedTime(&gms, e->ev0, e->ev1));
    st.gpu_ms += gms;
    st.alg_bytes += 4.0 * L.b.cols_read * (double)L.b.d.num_tuples;
    st.launches++;
    st.shared_queries += L.b.shared;
    // learn from the measured launch
    if (e->cm) {
      launch_features(L.b, L.es, L.b.d.num_tuples, L.nreset, cfg, feat);
      rc = mordred_cost_model_observe(e->cm, feat, std::max(1e-3, (double)gms), analytic_launch_ms(feat, L.b.d.ncols));
      if (rc) return fail(e, std::string("cost model: ") + mordred_cost_model_last_error(e->cm));
      if (++since_train >= 16) {
        since_train = 0;
        rc = mordred_cost_model_train(e->cm, 24, 128, 2e-3f, nullptr);
        if (rc) return fail(e, std::string("cost model: ") + mordred_cost_model_last_error(e->cm));
      }
      HIP_OK(hipSetDevice(e->device));
    }
    for (const auto& [sidx, req] : L.retire) {
      std::memcpy(res[req], stage_of(sidx), result_ints(req) * sizeof(int));
      times[req].done_ms = now_ms();
      done++;
    }
    return 0;
  };
  Launch fly;  // the launch on the GPU
  while (done < n) {
    // a lone request on an idle loop: alone through its own fused kernel
    if (!fly.on && o->lone_plain && active == 0 && next < n && arrival_ms[next] <= now_ms() &&
        (next + 1 >= n || arrival_ms[next + 1] > now_ms())) {
      const int i = next++;
      times[i].admit_ms = now_ms();
      mordred_run_stats rs{};
      const int rc = mordred_engine_run(e, &qs[i], MORDRED_PLACE_RESIDENCY, res[i], &rs);
      if (rc) return rc;
      if (rs.cpu_rows) return fail(e, "serve: fact columns must be cached in HBM");
      times[i].done_ms = now_ms();
      times[i].launches = 1;
      st.lone_requests++;  // (launches, gpu_ms and alg_bytes count shared-scan launches only)
      st.max_active_seen = std::max(st.max_active_seen, 1);
      done++;
      HIP_OK(hipSetDevice(e->device));
      continue;
    }
    // admission: free slot, arrived request, budget
    unsigned long long reset = 0;
    for (int sidx = 0; sidx < nslot && next < n && arrival_ms[next] <= now_ms(); sidx++) {
      if (slots[sidx].req >= 0) continue;
      if (o->hbm_budget_bytes > 0 && state + state_bytes(next) > o->hbm_budget_bytes && active > 0) break;
      slots[sidx] = Slot{next, cur, 0};
      qp[sidx] = &qs[next];
      times[next].admit_ms = now_ms();
      state += state_bytes(next);
      st.peak_state_bytes = std::max(st.peak_state_bytes, state);
      reset |= 1ull << sidx;
      active++;
      next++;
      changed = true;
    }
    Launch nx;
    if (active > 0) {
      st.max_active_seen = std::max(st.max_active_seen, active);
      // launch range: from cur, never past the table end or an active query's
      // starting segment (each query sees every segment exactly once)
      int limit = nseg;
      for (const Slot& sl : slots)
        if (sl.req >= 0 && sl.start > cur) limit = std::min(limit, sl.start);
      nx.es.assign(nslot, QueryEst{});
      std::vector<double> sel(nslot, 0.0);
      for (int sidx = 0; sidx < nslot; sidx++)
        if (slots[sidx].req >= 0) {
          nx.es[sidx] = est[slots[sidx].req];
          sel[sidx] = nx.es[sidx].sel;
        }
      BatchPlan& b = nx.b;
      int rc = plan_batch(e, qp.data(), nslot, cfg, b, &ar, sel.data());
      if (rc) return rc;
      b.d.reset = reset;
      b.d.rebuild = changed ? 1 : 0;
      for (int sidx = 0; sidx < nslot; sidx++) nx.nreset += reset >> sidx & 1ull;
      // launch size: the largest candidate predicted within the target
      int span = limit - cur;
      bool used_learned = false;
      if (o->chunk_segments > 0) {
        span = std::min(span, o->chunk_segments);
      } else {
        mordred_cost_model_stats cs{};
        const bool trusted = o->learned && e->cm && mordred_cost_model_get_stats(e->cm, &cs) == 0 && cs.trusted;
        std::vector<float> xs;
        std::vector<int> cand;
        for (int sz : sizes) {
          const int sp = std::min(sz, limit - cur);
          if (!cand.empty() && sp == cand.back()) break;
          cand.push_back(sp);
          launch_features(b, nx.es, rows_of(cur, cur + sp), nx.nreset, cfg, feat);
          xs.insert(xs.end(), feat, feat + MORDRED_COST_FEATURES);
        }
        std::vector<float> pred(cand.size());
        if (trusted) {
          rc = mordred_cost_model_predict(e->cm, xs.data(), (int)cand.size(), pred.data());
          if (rc) return fail(e, std::string("cost model: ") + mordred_cost_model_last_error(e->cm));
          HIP_OK(hipSetDevice(e->device));
          for (float& p : pred) p = std::exp(p);
          used_learned = true;
        } else {
          for (size_t c = 0; c < cand.size(); c++)
            pred[c] = (float)analytic_launch_ms(&xs[c * MORDRED_COST_FEATURES], b.d.ncols);
        }
        span = cand[0];
        for (size_t c = 0; c < cand.size(); c++)
          if (pred[c] <= o->target_chunk_ms) span = cand[c];
        (used_learned ? st.learned_choices : st.analytic_choices)++;
      }
      const int end = cur + span;
      b.d.seg_begin = cur;
      b.d.num_tuples = (int)rows_of(cur, end);
      changed = false;
      // account the launch now; queries that will have seen every segment
      // leave their slots to the next plan
      for (int sidx = 0; sidx < nslot; sidx++) {
        Slot& sl = slots[sidx];
        if (sl.req < 0) continue;
        sl.covered += span;
        times[sl.req].launches++;
        if (sl.covered < nseg) continue;
        nx.retire.emplace_back(sidx, sl.req);
        state -= state_bytes(sl.req);
        sl = Slot{};
        qp[sidx] = nullptr;
        active--;
        changed = true;
      }
      cur = end >= nseg ? 0 : end;
      nx.on = true;
    }
    if (fly.on) {
      const int rc = complete(fly);
      if (rc) return rc;
    }
    if (nx.on) {
      const int rc = launch(nx);
      if (rc) return rc;
      fly = std::move(nx);
      fly.on = true;
    } else if (!fly.on && active == 0 && next < n) {
      wait_until(arrival_ms[next]);
    }
  }
  st.wall_ms = now_ms();
  if (e->cm) {
    mordred_cost_model_stats cs{};
    if (mordred_cost_model_get_stats(e->cm, &cs) == 0) {
      st.learned_error = cs.ewma_abs_log_error;
      st.analytic_error = cs.ewma_abs_log_error_analytic;
      st.learned_trusted = cs.trusted;
    }
  }
  if (stats) *stats = st;
  return 0;
}

}  // extern "C"
