// pakhip_builders.hpp -- C++ adapters with the shapes of the reference's
// builder interfaces, over the C ABI of pakhip.h (libpakhip.so).
//
//   pakhip::Hip_Kmer_Partial_Counter<kmer_t>
//       drop-in for Concurrent_Kmer_Partial_Counter<kmer_t>
//       (src/concurrent_kmer_counter.hpp:21-68 ctor / reset, :2018-2082
//       default_count_and_process).  Same constructor arguments, same
//       return value (reads counted), same processor protocol
//       (processor(begin, end) over pairs with ->first = k-mer word,
//       ->second = int count), std::logic_error on failure.  Also
//       default_count_and_process_sort and concurrent_count_and_process_omp_v3
//       (:1832-1930, the one the reference's unit test calls), and the three
//       counters src/concurrent_kmer_table_builder.cpp:97-116 calls:
//       concurrent_count_and_process_omp_explicit (:519-666) and
//       _omp_implicit_v0 / _v1 (:910-1022, :1363-1578).  Those three never
//       call the processor: they return the reads and print the unique count
//       and the instance count as the reference does (at one thread).  The
//       remaining concurrent_* variants are left out: their results depend on
//       the OpenMP thread count (per-thread byte or window ranges).
//   pakhip::count_kmers / pakhip::kmer_histogram / pakhip::macro_nodes
//       the main path (perform_kmer_counting, src/kmer_counting.cpp:3351;
//       construct_kmer_histogram, src/kmer_histogram.cpp:72;
//       construct_macro_nodes, src/mn_construct.cpp:212) as value-returning
//       calls on one device context.
//   pakhip::visit_value / pakhip::reference_prefix_order
//       the per-edge visit count and the reference's prefix order, for
//       callers that fill the reference's MacroNode records.
//   pakhip::build_pakgraph_nodes
//       the node records of the PaK-graph builder (Pakman_Subgraph::
//       build_graph, src/pakman_graph.hpp:592-673) at one rank.
//
// Header-only; link with -lpakhip.  Host code only (no HIP headers needed).
#ifndef PAKHIP_BUILDERS_HPP
#define PAKHIP_BUILDERS_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pakhip.h"

namespace pakhip {

inline void throw_on_error(pk_ctx* ctx, int rc, const char* what) {
  if (rc == PK_OK) return;
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
  if (ctx) msg += ": " + std::string(pk_last_error(ctx));
  throw std::logic_error(msg);
}

// One device context (HIP stream + device memory); not copyable.
class Device_Context {
 public:
  explicit Device_Context(int device = 0, uint64_t hbm_budget_bytes = 0) {
    throw_on_error(nullptr, pk_ctx_create(&ctx_, device, hbm_budget_bytes), "pk_ctx_create");
  }
  ~Device_Context() { pk_ctx_destroy(ctx_); }
  Device_Context(const Device_Context&) = delete;
  Device_Context& operator=(const Device_Context&) = delete;
  pk_ctx* get() const { return ctx_; }

 private:
  pk_ctx* ctx_ = nullptr;
};

// kmer_t -> key words of the table (1: 64-bit word, 2: 128-bit word).
template <class kmer_t>
struct kmer_word_traits {
  static_assert(sizeof(kmer_t) == 8 || sizeof(kmer_t) == 16, "kmer_t must be a 64- or 128-bit word");
  static constexpr int bits = (int)sizeof(kmer_t) * 8;
  static kmer_t from_words(const uint64_t* w) {
    if (sizeof(kmer_t) == 8) return (kmer_t)w[0];
    return (kmer_t)(((unsigned __int128)w[1] << 64) | (unsigned __int128)w[0]);
  }
};

template <class kmer_t, template <class> class alloc = std::allocator>
class Hip_Kmer_Partial_Counter {
 public:
  typedef std::pair<kmer_t, int> kmer_pair_t;

  Hip_Kmer_Partial_Counter(size_t klen = 32, const char* symbol_begin = NULL,
                           const char* symbol_end = NULL, size_t read_len = 100,
                           size_t max_kmer_count = 1000000, char delimiter = '\n',
                           char skip_symbol = '>', int device = 0)
      : klen_(klen), symbol_begin_(symbol_begin), symbol_end_(symbol_end), read_len_(read_len),
        max_kmer_count_(max_kmer_count), delimiter_(delimiter), skip_symbol_(skip_symbol),
        ctx_(device) {}

  void reset(char const* begin, char const* end) {
    symbol_begin_ = begin;
    symbol_end_ = end;
  }

  // Counts on the device, then hands the ascending (k-mer, count) pairs to
  // the processor: one call when they number at most max_kmer_count (the
  // reference's trailing call; every reference caller passes SIZE_MAX),
  // otherwise consecutive slices of max_kmer_count pairs.  Returns the
  // number of reads counted.
  template <typename processor_t>
  size_t default_count_and_process(processor_t& processor) {
    std::vector<kmer_pair_t, alloc<kmer_pair_t>> pairs;
    const size_t reads = count_pairs(pairs);
    const size_t step = max_kmer_count_ ? max_kmer_count_ : 1;
    if (pairs.size() <= step) {
      processor(pairs.cbegin(), pairs.cend());
    } else {
      for (size_t b = 0; b < pairs.size(); b += step) {
        const size_t e = b + step < pairs.size() ? b + step : pairs.size();
        processor(pairs.cbegin() + b, pairs.cbegin() + e);
      }
    }
    symbol_begin_ = symbol_end_;  // the reference's scan consumes the range
    return reads;
  }

  // The reference's sibling counters produce the same table; on the device
  // they are one implementation.
  template <typename processor_t>
  size_t default_count_and_process_sort(processor_t& processor) {
    return default_count_and_process(processor);
  }

  // concurrent_count_and_process_omp_v3 (concurrent_kmer_counter.hpp:
  // 1832-1930, the variant the reference's unit test drives): the same
  // records, k-mers and read count as default_count_and_process on FASTA whose
  // '>' bytes start records, handed to the processor in one call over the
  // whole table (whatever max_kmer_count is), and the range is not consumed
  // (the reference scans it with local pointers).
  template <typename processor_t>
  size_t concurrent_count_and_process_omp_v3(processor_t& processor) {
    std::vector<kmer_pair_t, alloc<kmer_pair_t>> pairs;
    const size_t reads = count_pairs(pairs);
    processor(pairs.cbegin(), pairs.cend());
    return reads;
  }

  // concurrent_count_and_process_omp_explicit (concurrent_kmer_counter.hpp:
  // 519-666): the default counter's records and packed k-mers, sorted and
  // run-length counted; the processor is not called.  Prints the number of
  // runs ("unique kmers", the distinct k-mers; 1 when there is no k-mer, the
  // reference's one-thread figure, which pushes one frequency for an empty
  // range) and the instances ("concurrent_vector size").  Returns the reads.
  template <typename processor_t>
  size_t concurrent_count_and_process_omp_explicit(processor_t& /*processor*/) {
    pk_table t;
    uint64_t reads = 0;
    const size_t n = range_bytes();
    throw_on_error(ctx_.get(),
                   pk_builder_count(ctx_.get(), n ? symbol_begin_ : nullptr, n, (int)klen_, read_len_,
                                    kmer_word_traits<kmer_t>::bits, delimiter_, skip_symbol_, 0, &t, &reads),
                   "pk_builder_count");
    pk_timing tm;
    throw_on_error(ctx_.get(), pk_last_timing(ctx_.get(), &tm), "pk_last_timing");
    printf("unique kmers=%lu\n", (unsigned long)(tm.n_instances ? t.n : 1));
    printf("concurrent_vector size=%lu\n", (unsigned long)tm.n_instances);
    return (size_t)reads;
  }

  // concurrent_count_and_process_omp_implicit_v0 (concurrent_kmer_counter.hpp:
  // 910-1022): every record's klen-byte windows ordered by strncmp and
  // counted as distinct byte strings (no packing); a read cut short by the
  // end of the range throws std::logic_error("Variable size read").  The
  // processor is not called; prints as omp_explicit; returns the records.
  template <typename processor_t>
  size_t concurrent_count_and_process_omp_implicit_v0(processor_t& /*processor*/) {
    return count_windows(PK_WINDOWS_V0);
  }

  // concurrent_count_and_process_omp_implicit_v1 (concurrent_kmer_counter.hpp:
  // 1363-1578): as _v0, reads cut by the end of the range taken as they are
  // (a read shorter than klen throws here; the reference's count wraps).
  template <typename processor_t>
  size_t concurrent_count_and_process_omp_implicit_v1(processor_t& /*processor*/) {
    return count_windows(PK_WINDOWS_V1);
  }

  pk_ctx* context() const { return ctx_.get(); }

 private:
  size_t range_bytes() const {
    return (symbol_begin_ && symbol_end_ && symbol_end_ > symbol_begin_) ? (size_t)(symbol_end_ - symbol_begin_) : 0;
  }

  size_t count_windows(int mode) {
    if (klen_ > (size_t)std::numeric_limits<int>::max()) throw std::logic_error("klen too large");
    const size_t n = range_bytes();
    uint64_t reads = 0, windows = 0, distinct = 0;
    throw_on_error(ctx_.get(),
                   pk_builder_count_windows(ctx_.get(), n ? symbol_begin_ : nullptr, n, (int)klen_, read_len_,
                                            delimiter_, skip_symbol_, mode, 0, &reads, &windows, &distinct),
                   "pk_builder_count_windows");
    printf("unique kmers=%lu\n", (unsigned long)(windows ? distinct : 1));
    printf("concurrent_vector size=%lu\n", (unsigned long)windows);
    return (size_t)reads;
  }

  // the (k-mer, count) pairs of the counter's range, ascending; returns the
  // reads counted
  size_t count_pairs(std::vector<kmer_pair_t, alloc<kmer_pair_t>>& pairs) {
    const size_t n = range_bytes();
    if (klen_ > (size_t)std::numeric_limits<int>::max())
      throw std::logic_error("klen too large");
    pk_table t;
    uint64_t reads = 0;
    throw_on_error(ctx_.get(),
                   pk_builder_count(ctx_.get(), n ? symbol_begin_ : nullptr, n, (int)klen_,
                                    read_len_, kmer_word_traits<kmer_t>::bits, delimiter_,
                                    skip_symbol_, 0, &t, &reads),
                   "pk_builder_count");
    std::vector<uint64_t> words((size_t)t.n * t.words);
    std::vector<int32_t> counts((size_t)t.n);
    if (t.n)
      throw_on_error(ctx_.get(), pk_download_table(ctx_.get(), words.data(), counts.data()),
                     "pk_download_table");
    pairs.clear();
    pairs.reserve((size_t)t.n);
    for (size_t i = 0; i < (size_t)t.n; i++)
      pairs.emplace_back(kmer_word_traits<kmer_t>::from_words(&words[i * t.words]), counts[i]);
    return (size_t)reads;
  }

  size_t klen_;
  const char* symbol_begin_;
  const char* symbol_end_;
  size_t read_len_;
  size_t max_kmer_count_;
  char delimiter_;
  char skip_symbol_;
  Device_Context ctx_;
};

// -- main path -------------------------------------------------------------

struct Kmer_Table {
  int k = 0;
  int words = 1;
  std::vector<uint64_t> keys;  // n * words, lo word first
  std::vector<int32_t> counts;
  size_t size() const { return counts.size(); }
};

// perform_kmer_counting (src/kmer_counting.cpp:3351) on one device: the
// sorted unique (k-mer, count) table the reference leaves in kmer_proc_buf.
inline Kmer_Table count_kmers(pk_ctx* ctx, const char* read_data, size_t length, int k) {
  pk_table t;
  throw_on_error(ctx, pk_count(ctx, read_data, length, k, 0, &t), "pk_count");
  Kmer_Table out;
  out.k = t.k;
  out.words = t.words;
  out.keys.resize((size_t)t.n * t.words);
  out.counts.resize((size_t)t.n);
  if (t.n) throw_on_error(ctx, pk_download_table(ctx, out.keys.data(), out.counts.data()), "pk_download_table");
  return out;
}

struct Histogram_Cutoff {
  std::vector<uint64_t> hist;  // hist[c] for c < t
  int min_bucket = 0;
  Kmer_Table kept;
};

// construct_kmer_histogram (src/kmer_histogram.cpp:72-249, one rank) on the
// table of the last count_kmers.  num_buckets == 0 means 21.
inline Histogram_Cutoff kmer_histogram(pk_ctx* ctx, int num_buckets, int k) {
  Histogram_Cutoff r;
  const int t = num_buckets == 0 ? 21 : num_buckets;
  r.hist.assign((size_t)(t > 0 ? t : 1), 0);
  uint64_t nk = 0;
  throw_on_error(ctx, pk_cutoff(ctx, num_buckets, r.hist.data(), &r.min_bucket, &nk), "pk_cutoff");
  pk_table kt;
  throw_on_error(ctx, pk_kept_table(ctx, &kt), "pk_kept_table");
  r.kept.k = k;
  r.kept.words = kt.words;
  r.kept.keys.resize((size_t)kt.n * kt.words);
  r.kept.counts.resize((size_t)kt.n);
  if (kt.n) throw_on_error(ctx, pk_download_kept(ctx, r.kept.keys.data(), r.kept.counts.data()), "pk_download_kept");
  return r;
}

// construct_macro_nodes (src/mn_construct.cpp:212-477, one rank) on the kept
// table of the last kmer_histogram: node records ascending by (k-1)-mer key.
inline std::vector<pk_mnode> macro_nodes(pk_ctx* ctx) {
  pk_mn_table m;
  throw_on_error(ctx, pk_macro_nodes(ctx, &m), "pk_macro_nodes");
  std::vector<pk_mnode> nodes((size_t)m.n);
  if (m.n) throw_on_error(ctx, pk_download_nodes(ctx, nodes.data()), "pk_download_nodes");
  return nodes;
}

// visit_value (src/mn_construct.cpp:75-79): the visit count of an edge of
// `count` instances at `coverage`, computed as the reference does (double
// divide, ceil).
inline int visit_value(int count, int coverage) {
  return (int)std::ceil((double)count / (double)coverage);
}

// Reference prefix order (src/mn_construct.cpp:87-95,299-303,395-457): the
// reference walks a copy of the kept list sorted by the suffix key only
// (Comp_suffix), so a node's prefixes come in the order std::sort leaves
// tied elements.  This repeats that sort (same algorithm, same input order:
// the kept list ascending, same comparisons) and returns, for every node of
// `nodes` (ascending keys), its prefix bases in the reference's order (255 =
// unused slot).  Exact against a one-thread reference run built with the
// same libstdc++; the device's records carry the counts by base.
inline std::vector<std::array<uint8_t, 4>> reference_prefix_order(const Kmer_Table& kept,
                                                                  const std::vector<pk_mnode>& nodes) {
  typedef unsigned __int128 u128;
  const int k = kept.k;
  const size_t n = kept.size();
  std::vector<u128> seq(n);
  for (size_t i = 0; i < n; i++)
    seq[i] = kept.words == 1 ? (u128)kept.keys[i]
                             : (((u128)kept.keys[2 * i + 1] << 64) | (u128)kept.keys[2 * i]);
  const u128 succ_mask = (((u128)1) << (2 * (k - 1))) - 1;
  if (kept.words == 1) {
    std::vector<uint64_t> s64(n);
    for (size_t i = 0; i < n; i++) s64[i] = (uint64_t)seq[i];
    const uint64_t m = (uint64_t)succ_mask;
    std::sort(s64.begin(), s64.end(), [m](uint64_t a, uint64_t b) { return (a & m) < (b & m); });
    for (size_t i = 0; i < n; i++) seq[i] = s64[i];
  } else {
    std::sort(seq.begin(), seq.end(),
              [succ_mask](u128 a, u128 b) { return (a & succ_mask) < (b & succ_mask); });
  }
  std::vector<std::array<uint8_t, 4>> order(nodes.size());
  for (auto& o : order) o = {255, 255, 255, 255};
  size_t nd = 0;
  for (size_t i = 0; i < n; i++) {
    const u128 succ = seq[i] & succ_mask;
    while (nd < nodes.size() &&
           ((((u128)nodes[nd].key_hi) << 64) | (u128)nodes[nd].key_lo) < succ)
      nd++;
    if (nd == nodes.size()) break;
    const uint8_t base = (uint8_t)((seq[i] >> (2 * (k - 1))) & 3);
    for (int j = 0; j < 4; j++)
      if (order[nd][j] == 255) {
        order[nd][j] = base;
        break;
      }
  }
  return order;
}

// -- PaK-graph builder nodes ---------------------------------------------------
//
// Pakman_Subgraph::build_graph (src/pakman_graph.hpp:592-673) at one rank:
// k-mers with count > threshold, grouped by (k-1)-mer; every node holds a
// terminal prefix and suffix ("" , 0, 0) plus one single-base entry per kept
// neighbour (count, visit = ceil(count / coverage)); node_t::
// rebuild_prefix_suffix_connections (:357-386) then sorts both lists with
// visit_count_sorter_t (:80-108), gives the terminal of the lighter side the
// difference of the visit totals (:389-408) and wires prefixes to suffixes.
// The device supplies the kept table and the (k-1)-mer records
// (pk_cutoff_threshold + pk_macro_nodes); this assembles the node records.
struct Pakgraph_Entry {
  char base;  // 'A' 'C' 'T' 'G', or 0 for the terminal ("")
  int count;
  int visit;
  bool terminal;
  struct Wire {
    size_t suffix_idx;
    int count;
    int offset;
  };
  std::vector<Wire> wires;  // prefixes only
};

struct Pakgraph_Node {
  uint64_t key_lo, key_hi;  // (k-1)-mer
  std::vector<Pakgraph_Entry> prefixes, suffixes;
};

inline bool pakgraph_entry_before(const Pakgraph_Entry& a, const Pakgraph_Entry& b) {
  if (a.visit != b.visit) return a.visit > b.visit;
  if (a.count != b.count) return a.count > b.count;
  if (a.terminal != b.terminal) return !a.terminal > !b.terminal;
  // suffix_ordering: shorter string first, then symbol value A < C < T < G
  const int la = a.base ? 1 : 0, lb = b.base ? 1 : 0;
  if (la != lb) return la < lb;
  auto val = [](char c) { return c == 'A' ? 0 : c == 'C' ? 1 : c == 'T' ? 2 : 3; };
  return la ? val(a.base) < val(b.base) : false;
}

inline void pakgraph_rebuild(Pakgraph_Node& n) {
  std::sort(n.prefixes.begin(), n.prefixes.end(), pakgraph_entry_before);
  std::sort(n.suffixes.begin(), n.suffixes.end(), pakgraph_entry_before);
  int tp = 0, ts = 0;
  for (const auto& p : n.prefixes) tp += p.visit;
  for (const auto& s : n.suffixes) ts += s.visit;
  if (ts < tp) {
    n.suffixes.back().visit = tp - ts;
    n.suffixes.back().count = 1;
  } else if (ts > tp) {
    n.prefixes.back().visit = ts - tp;
    n.prefixes.back().count = 1;
  }
  size_t pi = 0, si = 0;
  int pu = 0, su = 0;
  while (pi < n.prefixes.size() && si < n.suffixes.size()) {
    const int delta = std::min(n.prefixes[pi].visit - pu, n.suffixes[si].visit - su);
    pu += delta;
    su += delta;
    n.prefixes[pi].wires.push_back({si, delta, su});
    if (pu == n.prefixes[pi].visit) {
      ++pi;
      pu = 0;
    }
    if (su == n.suffixes[si].visit) {
      ++si;
      su = 0;
    }
  }
}

// Nodes of the PaK-graph builder over the current table of ctx (after a
// count): threshold as Pakman_Subgraph::build_graph takes it (the harness
// passes t - 1), coverage as its `cov` argument.
inline std::vector<Pakgraph_Node> build_pakgraph_nodes(pk_ctx* ctx, uint64_t threshold, int coverage) {
  uint64_t nk = 0;
  throw_on_error(ctx, pk_cutoff_threshold(ctx, threshold, &nk), "pk_cutoff_threshold");
  std::vector<Pakgraph_Node> out;
  if (nk == 0) return out;
  const std::vector<pk_mnode> recs = macro_nodes(ctx);
  static const char kBase[4] = {'A', 'C', 'T', 'G'};
  out.resize(recs.size());
  for (size_t i = 0; i < recs.size(); i++) {
    Pakgraph_Node& n = out[i];
    n.key_lo = recs[i].key_lo;
    n.key_hi = recs[i].key_hi;
    n.prefixes.push_back({0, 0, 0, true, {}});
    n.suffixes.push_back({0, 0, 0, true, {}});
    for (int b = 0; b < 4; b++) {
      if (recs[i].pre_cnt[b])
        n.prefixes.push_back({kBase[b], recs[i].pre_cnt[b], visit_value(recs[i].pre_cnt[b], coverage), false, {}});
      if (recs[i].suf_cnt[b])
        n.suffixes.push_back({kBase[b], recs[i].suf_cnt[b], visit_value(recs[i].suf_cnt[b], coverage), false, {}});
    }
    pakgraph_rebuild(n);
  }
  return out;
}

}  // namespace pakhip

#endif  // PAKHIP_BUILDERS_HPP
