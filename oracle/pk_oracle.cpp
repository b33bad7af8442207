// pk_oracle.cpp -- CPU restatement of PaKman's k-mer counting, histogram
// cutoff and macro-node construction path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library
// is compared against; it is loaded only by tests/, by
// __graft_entry__.smoke() and by bench.py's CPU-baseline leg.  Nothing in the
// product (pakman_plus_amd/, include/) links, loads or calls it.
//
// Every function cites the reference file:line whose behaviour it restates
// (paths relative to the vkundeti/pakman_plus source tree).
//
// Parity pins (what checks this file, besides the HIP library):
//   * parse + pack + count: the brute-force checker pko_bruteforce_count()
//     below (closed-form per-window packing, std::map), the invariant
//     sum(counts) == sum_reads max(0, len-k+1), and the reference test's
//     golden invariant (2 reads, k=32 -> 138; src/concurrent_kmer_counter_test.cpp:96,113).
//   * histogram / min_bucket / kept count / macro nodes / prefix order:
//     the reference's own kmer_histogram.cpp + mn_construct.cpp + wire.cpp,
//     compiled from the reference tree into oracle/_ref by oracle/build_ref.sh
//     and run on the same tables (tests/golden/ fixtures).
//   * owner hash (rank-share mode): closed form of hash31 on small values and
//     the invariant that the P rank shares partition the P=1 table.
// Nothing here is "parity unpinned" except pko_builder_count's behaviour for
// max_kmer_count smaller than the number of distinct k-mers (see its comment).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <thread>
#include <utility>
#include <vector>

typedef unsigned __int128 u128;

namespace {

// ---------------------------------------------------------------------------
// Error codes: identical values to include/pakhip.h (PK_*).
enum {
  E_OK = 0,
  E_INVAL = -1,
  E_NOMEM = -2,
  E_FORMAT = -4,
  E_EMPTY = -5,
};

const int LMER_LENGTH = 8;                  // Makefile:2 -DLMER_LENGTH=8
const uint64_t LMER_MASK = (1ull << (2 * LMER_LENGTH)) - 1;  // distribute_kmers.h:97
const uint64_t HASHA = 68111, HASHB = 105929;  // distribute_kmers.cpp:89-90

// char_to_el: distribute_kmers.h:784-786.  A=0 C=1 T=2 G=3, N->7, '\r'->6.
inline unsigned char_to_el(char ch) { return (((unsigned char)ch) >> 1) & 0x7; }

template <class K>
struct Enc {
  int k;
  K kmask;   // KMER_MASK  distribute_kmers.h:95,101
  K mnmask;  // MN_MASK == SUCC_MASK  distribute_kmers.h:98-99,104-105
  explicit Enc(int k_) : k(k_) {
    const int bits = (int)sizeof(K) * 8;
    kmask = (~K(0)) >> (bits - 2 * k);
    mnmask = (~K(0)) >> (bits - 2 * (k - 1));
  }
  // kmer_shift: distribute_kmers.h:798-805 (el is the 3-bit code: bit 2 ORs
  // into the previous base's low bit).
  K shift(K km, unsigned el) const { return ((km << 2) | (K)el) & kmask; }
  // pred: mn_shift((kmer)^1) kmer_histogram.cpp:150-151, mn_construct.cpp:234-235
  K pred(K km) const { return ((km ^ (K)1) >> 2) & mnmask; }
  // succ: kmer & SUCC_MASK  kmer_histogram.cpp:154
  K succ(K km) const { return km & mnmask; }
  // kmerel: distribute_kmers.h:755-759
  unsigned el(K km, int pos) const { return (unsigned)((km >> ((k - pos - 1) * 2)) & 3); }
};

// hash31: distribute_kmers.h:856-873 (signed 64-bit, arithmetic shift).
inline long hash31(long long a, long long b, long long x) {
  long long result = (long long)((unsigned long long)a * (unsigned long long)x +
                                 (unsigned long long)b);
  result = ((result >> 31) + result) & 2147483647LL;
  return (long)result;
}
// retrieve_proc_id(lmer_t): distribute_kmers.cpp:95-105
inline int owner_of_lmer(uint64_t lmer, int P) {
  return (int)((int)hash31((long long)HASHA, (long long)HASHB, (long long)lmer) % P);
}

// kmer_to_lmer: distribute_kmers.h:774-782
template <class K>
inline uint64_t kmer_to_lmer(const Enc<K>& e, K kin, int pos, uint64_t lout) {
  return ((lout << 2) | (uint64_t)e.el(kin, pos)) & LMER_MASK;
}

// recalculate_min_lmer: kmer_counting.cpp:355-388 (leftmost strict minimum of
// the L-mer frequencies extracted from the packed k-mer).
template <class K>
void recalc_min_lmer(const Enc<K>& e, K kin, const std::vector<uint64_t>& freq,
                     uint64_t* m_lmer, uint64_t* m_freq, int* m_pos) {
  uint64_t min_lmer = 0, min_f = 0;
  int min_pos = 0, kk = 0;
  for (kk = 0; ((e.k - 1) - kk) >= (LMER_LENGTH - 1); kk++) {
    uint64_t lo = 0;
    for (int j = kk; j < LMER_LENGTH + kk; j++) lo = kmer_to_lmer(e, kin, j, lo);
    uint64_t f = freq[lo];
    if (kk == 0) {
      min_lmer = lo; min_f = f; min_pos = 0;
    } else if (f < min_f) {
      min_lmer = lo; min_f = f; min_pos = kk;
    }
  }
  *m_lmer = min_lmer; *m_freq = min_f; *m_pos = min_pos;
}

// Sliding_window_l: kmer_counting.cpp:391-438.  Counts every L-mer of every
// read line.  Returns E_FORMAT where the reference's assert would fire.
int lmer_frequency(const char* ptr, size_t length, std::vector<uint64_t>& freq) {
  freq.assign((size_t)1 << (2 * LMER_LENGTH), 0);
  size_t p = 0;
  for (; p < length && ptr[p] != '>'; p++) {}
  uint64_t km = 0;
  while (p < length) {
    if (ptr[p] != '>') return E_FORMAT;
    for (; p < length && ptr[p] != '\n'; p++) {}
    p++;
    if (p + LMER_LENGTH > length) break;
    km = 0;
    for (int i = 0; ptr[p] != '\n' && i < LMER_LENGTH - 1; i++)
      km = ((km << 2) | char_to_el(ptr[p++])) & LMER_MASK;
    while (p < length && ptr[p] != '\n') {
      km = ((km << 2) | char_to_el(ptr[p++])) & LMER_MASK;
      freq[km]++;
    }
    p++;
  }
  return E_OK;
}

// Sliding_window: kmer_counting.cpp:2890-3154.  Calls sink(kmer, owner) for
// every k-mer instance in input order.  With P == 1 the min-L-mer tracking is
// skipped (every owner is 0).  Returns E_FORMAT where the reference asserts
// (ptr[p] != '>' at a record start, kmer_counting.cpp:2928).
template <class K, class Sink>
int sliding_window(const char* ptr, size_t length, const Enc<K>& e, int P,
                   const std::vector<uint64_t>* gfreq, Sink&& sink) {
  const int KL = e.k;
  size_t p = 0;
  for (; p < length && ptr[p] != '>'; p++) {}
  while (p < length) {
    if (ptr[p] != '>') return E_FORMAT;
    for (; p < length && ptr[p] != '\n'; p++) {}
    p++;
    if (p + KL > length) break;  // too short a read (kmer_counting.cpp:2936)
    K km = 0;
    uint64_t lmer_out = 0, min_lmer = 0, min_f = 0;
    int min_pos = 0;
    int i;
    for (i = 0; ptr[p] != '\n' && i < KL - 1; i++) {
      unsigned el = char_to_el(ptr[p]);
      km = e.shift(km, el);
      if (P > 1) {
        lmer_out = ((lmer_out << 2) | el) & LMER_MASK;
        if (i >= LMER_LENGTH - 1) {
          uint64_t f = (*gfreq)[lmer_out];
          int tpos = i - (LMER_LENGTH - 1);
          if (i == LMER_LENGTH - 1) {
            min_lmer = lmer_out; min_f = f;
          } else if (f < min_f) {
            min_lmer = lmer_out; min_f = f; min_pos = tpos;
          }
        }
      }
      p++;
    }
    while (p < length && ptr[p] != '\n') {
      unsigned el = char_to_el(ptr[p]);
      km = e.shift(km, el);
      int owner = 0;
      if (P > 1) {
        lmer_out = ((lmer_out << 2) | el) & LMER_MASK;
        uint64_t f = (*gfreq)[lmer_out];
        if (min_pos < 0) recalc_min_lmer(e, km, *gfreq, &min_lmer, &min_f, &min_pos);
        if (f < min_f) {
          min_lmer = lmer_out; min_f = f; min_pos = KL - LMER_LENGTH;
        }
        min_pos--;
        owner = owner_of_lmer(min_lmer, P);
      }
      sink(km, owner);
      p++;
    }
    p++;
  }
  return E_OK;
}

template <class K>
struct Pair {
  K seq;
  int cnt;
};

// SortAndAggregate_InPlace: kmer_counting.cpp:169-192
template <class K>
void aggregate_inplace(std::vector<Pair<K>>& arr) {
  if (arr.empty()) return;
  size_t u = 0;
  int cum = arr[0].cnt;
  for (size_t i = 1; i < arr.size(); i++) {
    if (arr[i].seq != arr[u].seq) {
      arr[u].cnt = cum;
      ++u;
      arr[u].seq = arr[i].seq;
      cum = 0;
    }
    cum += arr[i].cnt;
  }
  arr[u].cnt = cum;
  arr.resize(u + 1);
}

// SortAndAggregate (lower/upper_bound form): kmer_counting.cpp:194-250
template <class K>
void aggregate_bounds(std::vector<Pair<K>>& arr) {
  std::vector<Pair<K>> out;
  auto lt = [](const Pair<K>& a, const K& b) { return a.seq < b; };
  auto gt = [](const K& b, const Pair<K>& a) { return b < a.seq; };
  for (auto it = arr.begin(); it != arr.end();) {
    K key = it->seq;
    auto lo = std::lower_bound(arr.begin(), arr.end(), key, lt);
    auto up = std::upper_bound(arr.begin(), arr.end(), key, gt);
    int sum = 0;
    for (auto q = lo; q != up; ++q) sum += q->cnt;
    out.push_back(Pair<K>{key, sum});
    it = up;
  }
  arr.swap(out);
}

// One batch flush for the destination this process keeps: per-destination
// std::sort + run-length count (kmer_counting.cpp:3057-3086, 3166-3196), then
// transfer_kmers' receive side (kmer_counting.cpp:300-337): sort received
// pairs by seq, append + inplace_merge into the table, aggregate.
template <class K>
void flush_batch(std::vector<K>& part, std::vector<Pair<K>>& table, bool inplace) {
  if (part.empty()) return;
  std::vector<Pair<K>> send;
  std::sort(part.begin(), part.end());
  K prev = part[0];
  int counter = 1;
  for (size_t i = 1; i < part.size(); i++) {
    if (part[i] == prev) {
      counter++;
    } else {
      send.push_back(Pair<K>{prev, counter});
      counter = 1;
      prev = part[i];
    }
  }
  send.push_back(Pair<K>{prev, counter});
  part.clear();
  part.shrink_to_fit();
  auto by_seq = [](const Pair<K>& a, const Pair<K>& b) { return a.seq < b.seq; };
  std::sort(send.begin(), send.end(), by_seq);
  size_t off = table.size();
  table.insert(table.end(), send.begin(), send.end());
  if (off) std::inplace_merge(table.begin(), table.begin() + off, table.end(), by_seq);
  if (inplace) aggregate_inplace(table); else aggregate_bounds(table);
}

// perform_kmer_counting's default path at P ranks, keeping rank r's share:
// kmer_counting.cpp:3351-3458 (Sliding_window_l + Allreduce, Sliding_window
// with -b batch flushes, process_remaining_kmers).  With P == 1 the result is
// the whole input's table.
template <class K>
int count_main(const char* fasta, size_t len, int k, int P, int r, uint64_t batch,
               bool inplace, std::vector<Pair<K>>& table) {
  Enc<K> e(k);
  std::vector<uint64_t> freq;
  if (P > 1) {
    int rc = lmer_frequency(fasta, len, freq);
    if (rc) return rc;
  }
  std::vector<K> part;
  uint64_t num = 0;
  table.clear();
  int rc = sliding_window<K>(fasta, len, e, P, P > 1 ? &freq : nullptr,
                             [&](K km, int owner) {
                               if (owner == r) part.push_back(km);
                               num++;
                               if (num > batch) {
                                 flush_batch(part, table, inplace);
                                 num = 0;
                               }
                             });
  if (rc) return rc;
  if (num) flush_batch(part, table, inplace);
  return E_OK;
}

// Brute force (independent of the functions above): split the buffer into
// records by a forward scan, evaluate every window with the closed form
//   kmer_i = (OR_j code3(c[i+j]) << 2(k-1-j)) & KMER_MASK
// and count with std::map.  Assumes a well-formed strict single-line FASTA.
template <class K>
void bruteforce(const char* s, size_t n, int k, std::map<K, long long>& out) {
  Enc<K> e(k);
  size_t i = 0;
  while (i < n && s[i] != '>') i++;
  while (i < n) {
    const char* nl = (const char*)memchr(s + i, '\n', n - i);
    if (!nl) break;
    size_t b = (size_t)(nl - s) + 1;
    if (b + (size_t)k > n) break;
    const char* nl2 = (const char*)memchr(s + b, '\n', n - b);
    size_t end = nl2 ? (size_t)(nl2 - s) : n;
    for (size_t w = b; w + (size_t)k <= end; w++) {
      K v = 0;
      for (int j = 0; j < k; j++) v |= (K)char_to_el(s[w + j]) << (2 * (k - 1 - j));
      out[v & e.kmask]++;
    }
    i = end + 1;
  }
}

// Builder-API counting: Concurrent_Kmer_Partial_Counter::default_count_and_process
// (concurrent_kmer_counter.hpp:2018-2082).  Fixed read_len stride after each
// header line, symbol_to_kmer (non-ACGT -> 0, :2097-2110), no KMER_MASK (the
// word keeps the last sizeof(K)*4 bases).  Returns the number of reads
// counted.  The reference hands partial maps to the processor whenever the
// map exceeds max_kmer_count; this restatement returns the aggregate of all
// calls, which is what an aggregating processor sees (parity of the partial
// call boundaries is unpinned and not reproduced).
template <class K>
size_t builder_count(const char* beg, const char* end, size_t klen, size_t read_len, char delim,
                     char skip, std::map<K, long long>& out) {
  auto sym = [](char c) -> K {
    switch (c) {
      case 'A': return 0;
      case 'C': return 1;
      case 'T': return 2;
      case 'G': return 3;
      default: return 0;
    }
  };
  size_t rid = 0;
  const char* cur = beg;
  while (cur != end) {
    while (cur != end && *cur != skip) ++cur;
    if (cur == end) break;
    ++cur;
    while (cur != end && *cur != delim) ++cur;
    if (cur == end) break;
    ++cur;
    size_t rl = read_len;
    const char* rb = cur;
    if ((size_t)(end - rb) < rl) rl = (size_t)(end - rb);
    if (rl < klen) continue;
    const char* wend = rb + (rl - klen + 1);
    K km = 0;
    const char* wk = rb;
    for (size_t i = 0; i < klen - 1; i++, wk++) km = (km << 2) | sym(*wk);
    for (const char* wb = rb; wb < wend; ++wb, ++wk) {
      km = (km << 2) | sym(*wk);
      out[km]++;
    }
    ++rid;
  }
  return rid;
}

// construct_kmer_histogram: kmer_histogram.cpp:84-140 (P = 1).
// hist[c] for 0 <= c < t; min_bucket = first argmin over hist[1..t-1]
// (std::min_element; an empty range gives t, i.e. 1 when t == 1); keep
// count >= min_bucket.
int cutoff(const int32_t* counts, size_t n, int t, uint64_t* hist, int* min_bucket,
           uint8_t* keep, size_t* n_kept) {
  if (t < 1) return E_INVAL;
  for (int i = 0; i < t; i++) hist[i] = 0;
  for (size_t f = 0; f < n; f++) {
    int v = counts[f];
    if (v >= 0 && v < t) hist[v] += 1;
  }
  int mb = t;
  if (t > 1) {
    mb = 1;
    for (int i = 2; i < t; i++)
      if (hist[i] < hist[mb]) mb = i;
  }
  size_t kept = 0;
  for (size_t f = 0; f < n; f++) {
    uint8_t kf = counts[f] >= mb;
    if (keep) keep[f] = kf;
    kept += kf;
  }
  *min_bucket = mb;
  *n_kept = kept;
  return kept ? E_OK : E_EMPTY;  // kmer_histogram.cpp:189-200 aborts on zero
}

// One macro node as produced by construct_macro_nodes.
template <class K>
struct Node {
  K key;
  int32_t pre_cnt[4];
  int32_t suf_cnt[4];
  uint8_t pre_order[4];  // prefix bases in list order (0xff = unused)
  uint8_t suf_order[4];
  uint8_t n_pre, n_suf;
};

// construct_macro_nodes at P = 1: mn_construct.cpp:212-477.
//  * node keys: sort + unique of pred and succ of every kept k-mer (:220-285)
//  * suffixes: walk the seq-sorted list, group by pred (:317-391)
//  * prefixes: walk a copy sorted with Comp_suffix (succ key only, :87-95,
//    :302-303), group by succ (:395-457).  ref_order != 0 reproduces the
//    single-thread reference order (std::sort with that comparator, which
//    is what __gnu_parallel::sort runs with one OpenMP thread); ref_order == 0
//    emits prefixes in ascending base order.
template <class K>
int macro_nodes(const Pair<K>* kept, size_t V, int k, int ref_order, std::vector<Node<K>>& nodes) {
  Enc<K> e(k);
  std::vector<K> keys;
  keys.reserve(2 * V);
  for (size_t f = 0; f < V; f++) {
    keys.push_back(e.pred(kept[f].seq));
    keys.push_back(e.succ(kept[f].seq));
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  nodes.assign(keys.size(), Node<K>());
  for (size_t i = 0; i < keys.size(); i++) {
    Node<K>& nd = nodes[i];
    nd.key = keys[i];
    for (int b = 0; b < 4; b++) {
      nd.pre_cnt[b] = nd.suf_cnt[b] = 0;
      nd.pre_order[b] = nd.suf_order[b] = 0xff;
    }
    nd.n_pre = nd.n_suf = 0;
  }
  std::vector<Pair<K>> by_seq(kept, kept + V);
  auto pair_cmp = [](const Pair<K>& a, const Pair<K>& b) { return a.seq < b.seq; };
  std::sort(by_seq.begin(), by_seq.end(), pair_cmp);
  std::vector<Pair<K>> by_suf = by_seq;
  const K m = e.mnmask;
  auto comp_suffix = [m](const Pair<K>& a, const Pair<K>& b) { return (a.seq & m) < (b.seq & m); };
  if (ref_order) {
    std::sort(by_suf.begin(), by_suf.end(), comp_suffix);
  } else {
    std::stable_sort(by_suf.begin(), by_suf.end(), comp_suffix);  // -> ascending first base
  }
  size_t lo = 0;
  for (size_t f = 0; f < V; f++) {
    K pk = e.pred(by_seq[f].seq);
    while (nodes[lo].key != pk) lo++;
    Node<K>& nd = nodes[lo];
    unsigned b = e.el(by_seq[f].seq, k - 1);  // kmerel(full_kmer, MN_LENGTH)
    nd.suf_cnt[b] = by_seq[f].cnt;
    nd.suf_order[nd.n_suf++] = (uint8_t)b;
  }
  lo = 0;
  for (size_t f = 0; f < V; f++) {
    K sk = e.succ(by_suf[f].seq);
    while (nodes[lo].key != sk) lo++;
    Node<K>& nd = nodes[lo];
    unsigned a = e.el(by_suf[f].seq, 0);  // kmerel(full_kmer, 0)
    nd.pre_cnt[a] = by_suf[f].cnt;
    nd.pre_order[nd.n_pre++] = (uint8_t)a;
  }
  return E_OK;
}

// ---------------------------------------------------------------------------
// Result containers handed across the C ABI.
struct Table {
  int words = 1;
  std::vector<uint64_t> keys;  // n * words (lo, hi)
  std::vector<int32_t> counts;
  uint64_t reads = 0;
};

template <class K>
inline void put_key(std::vector<uint64_t>& v, K x) {
  v.push_back((uint64_t)x);
  if (sizeof(K) > 8) v.push_back((uint64_t)((u128)x >> 64));
}
template <class K>
inline K get_key(const uint64_t* w, size_t i, int words) {
  if (words == 1) return (K)w[i];
  return (K)(((u128)w[2 * i + 1] << 64) | (u128)w[2 * i]);
}

struct Nodes {
  int words = 1;
  std::vector<uint64_t> keys;
  std::vector<int32_t> pre_cnt, suf_cnt;  // n * 4
  std::vector<uint8_t> pre_order, suf_order;  // n * 4
  std::vector<uint8_t> n_pre, n_suf;
};

bool valid_k(int k) { return k >= 8 && k <= 64; }

// ---------------------------------------------------------------------------
// The reference at P ranks, one host thread per rank (the CPU baseline of
// the benchmark: mpiexec -np P with OMP_NUM_THREADS=1, as the reference's own
// runs, scripts/cori/human_genome/run_human_str.sh:20,41).  Every phase is a
// parallel step over the ranks; joining the threads stands for the
// collective between phases.
//   1. DivideReads (src/input_reading.cpp:69-243): rank r reads from the first
//      '>' at or after r * (n / P) to the first '>' at or after (r+1) * (n / P).
//   2. Sliding_window_l on the share, summed over ranks (MPI_Allreduce,
//      src/kmer_counting.cpp:391-438, 3366).
//   3. Sliding_window (:2890-3154): each window goes to rank hash31(min L-mer)
//      % P; per destination std::sort + run-length count (:3057-3086,
//      3156-3261; one flush per rank).
//   4. transfer_kmers (:254-352): rank r sorts the pairs it receives and
//      aggregates them into its table.
//   5. construct_kmer_histogram (src/kmer_histogram.cpp:72-249): histogram
//      summed over ranks, min_bucket, each kept k-mer sent to the owners of
//      its pred and succ (retrieve_proc_id of kmer_t: hash31 for 64-bit keys,
//      uhash31 for 128-bit).
//   6. construct_macro_nodes (src/mn_construct.cpp:212-477): rank r builds the
//      nodes whose keys it owns from what it received.
// Totals: distinct k-mers, kept k-mers and nodes over the ranks (equal to the
// one-rank figures; tested), and the histogram.

inline long long uhash31_128(u128 x) {
  u128 r = (u128)HASHA * x + (u128)HASHB;
  r = ((r >> 31) + r) & (u128)2147483647;
  return (long long)r;
}
inline int node_owner(uint64_t x, int P) { return (int)((int)hash31((long long)HASHA, (long long)HASHB, (long long)x) % P); }
inline int node_owner(u128 x, int P) { return (int)((int)uhash31_128(x) % P); }

template <class F>
void for_ranks(int P, F&& f) {
  std::vector<std::thread> th;
  for (int r = 1; r < P; r++) th.emplace_back(f, r);
  f(0);
  for (auto& t : th) t.join();
}

struct RankRun {
  uint64_t distinct = 0, kept = 0, nodes = 0, instances = 0;
  int min_bucket = 0;
  std::vector<uint64_t> hist;
};

template <class K>
int pipeline_ranks(const char* fa, size_t len, int k, int t, int P, RankRun& out) {
  Enc<K> e(k);
  // 1. shares
  std::vector<size_t> lo(P), hi(P);
  const size_t step = len / (size_t)P;
  auto snap = [&](size_t pos) {
    while (pos < len && fa[pos] != '>') pos++;
    return pos;
  };
  for (int r = 0; r < P; r++) {
    lo[r] = r == 0 ? 0 : snap((size_t)r * step);
    hi[r] = r == P - 1 ? len : snap((size_t)(r + 1) * step);
    if (hi[r] < lo[r]) hi[r] = lo[r];
  }
  // 2. L-mer frequencies
  std::vector<std::vector<uint64_t>> lf(P);
  std::vector<int> rcs(P, E_OK);
  for_ranks(P, [&](int r) { rcs[r] = lmer_frequency(fa + lo[r], hi[r] - lo[r], lf[r]); });
  for (int r = 0; r < P; r++)
    if (rcs[r]) return rcs[r];
  std::vector<uint64_t> freq(lf[0].size(), 0);
  for (int r = 0; r < P; r++)
    for (size_t i = 0; i < freq.size(); i++) freq[i] += lf[r][i];
  // 3. route, sort and run-length count per destination
  std::vector<std::vector<std::vector<Pair<K>>>> sent(P, std::vector<std::vector<Pair<K>>>(P));
  std::vector<uint64_t> inst(P, 0);
  for_ranks(P, [&](int r) {
    std::vector<std::vector<K>> part(P);
    rcs[r] = sliding_window<K>(fa + lo[r], hi[r] - lo[r], e, P, P > 1 ? &freq : nullptr,
                               [&](K km, int owner) { part[owner].push_back(km); inst[r]++; });
    for (int d = 0; d < P; d++) {
      auto& v = part[d];
      std::sort(v.begin(), v.end());
      auto& o = sent[r][d];
      for (size_t i = 0; i < v.size();) {
        size_t j = i;
        while (j < v.size() && v[j] == v[i]) j++;
        o.push_back(Pair<K>{v[i], (int)(j - i)});
        i = j;
      }
      std::vector<K>().swap(v);
    }
  });
  for (int r = 0; r < P; r++)
    if (rcs[r]) return rcs[r];
  // 4. receive, sort, aggregate
  std::vector<std::vector<Pair<K>>> table(P);
  auto by_seq = [](const Pair<K>& a, const Pair<K>& b) { return a.seq < b.seq; };
  for_ranks(P, [&](int r) {
    auto& tb = table[r];
    for (int s = 0; s < P; s++) {
      tb.insert(tb.end(), sent[s][r].begin(), sent[s][r].end());
      std::vector<Pair<K>>().swap(sent[s][r]);
    }
    std::sort(tb.begin(), tb.end(), by_seq);
    aggregate_inplace(tb);
  });
  // 5. histogram, min_bucket, kept k-mers to their node owners
  out.hist.assign((size_t)t, 0);
  std::vector<std::vector<uint64_t>> h(P, std::vector<uint64_t>((size_t)t, 0));
  for_ranks(P, [&](int r) {
    for (auto& pr : table[r])
      if (pr.cnt >= 0 && pr.cnt < t) h[r][pr.cnt]++;
  });
  for (int r = 0; r < P; r++)
    for (int c = 0; c < t; c++) out.hist[c] += h[r][c];
  int mb = t;
  if (t > 1) {
    mb = 1;
    for (int i = 2; i < t; i++)
      if (out.hist[i] < out.hist[mb]) mb = i;
  }
  out.min_bucket = mb;
  std::vector<std::vector<std::vector<Pair<K>>>> mn(P, std::vector<std::vector<Pair<K>>>(P));
  std::vector<uint64_t> kept(P, 0);
  for_ranks(P, [&](int r) {
    for (auto& pr : table[r]) {
      if (pr.cnt < mb) continue;
      kept[r]++;
      const int hp = node_owner(e.pred(pr.seq), P), hs = node_owner(e.succ(pr.seq), P);
      mn[r][hp].push_back(pr);
      if (hs != hp) mn[r][hs].push_back(pr);
    }
  });
  // 6. macro nodes of the owned keys
  std::vector<uint64_t> nn(P, 0);
  for_ranks(P, [&](int r) {
    std::vector<Pair<K>> recv;
    for (int s = 0; s < P; s++) recv.insert(recv.end(), mn[s][r].begin(), mn[s][r].end());
    std::vector<Node<K>> nodes;
    if (!recv.empty()) macro_nodes<K>(recv.data(), recv.size(), k, 1, nodes);
    for (auto& nd : nodes)
      if (node_owner(nd.key, P) == r) nn[r]++;
  });
  for (int r = 0; r < P; r++) {
    out.distinct += table[r].size();
    out.kept += kept[r];
    out.nodes += nn[r];
    out.instances += inst[r];
  }
  return E_OK;
}

}  // namespace

extern "C" {

typedef struct pko_table pko_table;
typedef struct pko_nodes pko_nodes;

// Count k-mers of a FASTA buffer (main path).  nranks/rank select one
// rank's share (nranks == 1: whole input).  batch = -b MAX_KMER_COUNT.
int pko_count(const char* fasta, size_t len, int k, int nranks, int rank, uint64_t batch,
              int inplace, pko_table** out) {
  *out = nullptr;
  if (!valid_k(k) || nranks < 1 || rank < 0 || rank >= nranks || batch == 0) return E_INVAL;
  Table* t = new (std::nothrow) Table();
  if (!t) return E_NOMEM;
  int rc;
  if (k <= 32) {
    std::vector<Pair<uint64_t>> tab;
    rc = count_main<uint64_t>(fasta, len, k, nranks, rank, batch, inplace != 0, tab);
    t->words = 1;
    for (auto& p : tab) { put_key(t->keys, p.seq); t->counts.push_back(p.cnt); }
  } else {
    std::vector<Pair<u128>> tab;
    rc = count_main<u128>(fasta, len, k, nranks, rank, batch, inplace != 0, tab);
    t->words = 2;
    for (auto& p : tab) { put_key(t->keys, p.seq); t->counts.push_back(p.cnt); }
  }
  if (rc) { delete t; return rc; }
  *out = (pko_table*)t;
  return E_OK;
}

// Brute-force table (std::map over closed-form windows).
int pko_bruteforce_count(const char* fasta, size_t len, int k, pko_table** out) {
  *out = nullptr;
  if (!valid_k(k)) return E_INVAL;
  Table* t = new Table();
  if (k <= 32) {
    std::map<uint64_t, long long> m;
    bruteforce<uint64_t>(fasta, len, k, m);
    for (auto& kv : m) { put_key(t->keys, kv.first); t->counts.push_back((int32_t)kv.second); }
  } else {
    std::map<u128, long long> m;
    bruteforce<u128>(fasta, len, k, m);
    t->words = 2;
    for (auto& kv : m) { put_key(t->keys, kv.first); t->counts.push_back((int32_t)kv.second); }
  }
  *out = (pko_table*)t;
  return E_OK;
}

// Builder-API table (default_count_and_process).  kmer_bits = 64 or 128 is
// the width of the counter's kmer_t template argument.
int pko_builder_count(const char* fasta, size_t len, int klen, uint64_t read_len, int kmer_bits,
                      char delim, char skip, pko_table** out) {
  *out = nullptr;
  if (klen < 1 || (kmer_bits != 64 && kmer_bits != 128)) return E_INVAL;
  Table* t = new Table();
  if (kmer_bits == 64) {
    std::map<uint64_t, long long> m;
    t->reads = builder_count<uint64_t>(fasta, fasta + len, (size_t)klen, read_len, delim, skip, m);
    for (auto& kv : m) { put_key(t->keys, kv.first); t->counts.push_back((int32_t)kv.second); }
  } else {
    std::map<u128, long long> m;
    t->reads = builder_count<u128>(fasta, fasta + len, (size_t)klen, read_len, delim, skip, m);
    t->words = 2;
    for (auto& kv : m) { put_key(t->keys, kv.first); t->counts.push_back((int32_t)kv.second); }
  }
  *out = (pko_table*)t;
  return E_OK;
}

size_t pko_table_size(const pko_table* t) { return ((const Table*)t)->counts.size(); }
int pko_table_words(const pko_table* t) { return ((const Table*)t)->words; }
uint64_t pko_table_reads(const pko_table* t) { return ((const Table*)t)->reads; }
void pko_table_copy(const pko_table* tp, uint64_t* keys, int32_t* counts) {
  const Table* t = (const Table*)tp;
  if (keys && !t->keys.empty()) memcpy(keys, t->keys.data(), t->keys.size() * 8);
  if (counts && !t->counts.empty()) memcpy(counts, t->counts.data(), t->counts.size() * 4);
}
void pko_table_free(pko_table* t) { delete (Table*)t; }

// L-mer frequencies (65536 bins), Sliding_window_l.
int pko_lmer_frequency(const char* fasta, size_t len, uint64_t* freq_out) {
  std::vector<uint64_t> f;
  int rc = lmer_frequency(fasta, len, f);
  if (rc) return rc;
  memcpy(freq_out, f.data(), f.size() * 8);
  return E_OK;
}

// Owner of an L-mer at P ranks (retrieve_proc_id on lmer_t).
int pko_lmer_owner(uint64_t lmer, int nranks) { return owner_of_lmer(lmer, nranks); }

int pko_cutoff(const int32_t* counts, size_t n, int t, uint64_t* hist, int* min_bucket,
               uint8_t* keep, size_t* n_kept) {
  return cutoff(counts, n, t, hist, min_bucket, keep, n_kept);
}

// visit_value: mn_construct.cpp:75-79
int pko_visit_value(int num, int coverage) {
  double v = (double)num / coverage;
  return (int)ceil(v);
}

int pko_macro_nodes(const uint64_t* keys, const int32_t* counts, size_t V, int k, int ref_order,
                    pko_nodes** out) {
  *out = nullptr;
  if (!valid_k(k)) return E_INVAL;
  Nodes* nd = new Nodes();
  auto emit = [&](auto& nodes) {
    for (auto& x : nodes) {
      put_key(nd->keys, x.key);
      for (int b = 0; b < 4; b++) {
        nd->pre_cnt.push_back(x.pre_cnt[b]);
        nd->suf_cnt.push_back(x.suf_cnt[b]);
        nd->pre_order.push_back(x.pre_order[b]);
        nd->suf_order.push_back(x.suf_order[b]);
      }
      nd->n_pre.push_back(x.n_pre);
      nd->n_suf.push_back(x.n_suf);
    }
  };
  if (k <= 32) {
    std::vector<Pair<uint64_t>> kept(V);
    for (size_t i = 0; i < V; i++) kept[i] = Pair<uint64_t>{keys[i], counts[i]};
    std::vector<Node<uint64_t>> nodes;
    macro_nodes<uint64_t>(kept.data(), V, k, ref_order, nodes);
    nd->words = 1;
    emit(nodes);
  } else {
    std::vector<Pair<u128>> kept(V);
    for (size_t i = 0; i < V; i++) kept[i] = Pair<u128>{get_key<u128>(keys, i, 2), counts[i]};
    std::vector<Node<u128>> nodes;
    macro_nodes<u128>(kept.data(), V, k, ref_order, nodes);
    nd->words = 2;
    emit(nodes);
  }
  *out = (pko_nodes*)nd;
  return E_OK;
}

// The reference pipeline (count, cutoff, macro nodes) at nranks ranks, one
// host thread per rank (pipeline_ranks above).  stats: distinct, kept, nodes,
// instances, min_bucket; hist: t entries.  Returns E_EMPTY when nothing is
// kept (the reference aborts there).
int pko_pipeline_ranks(const char* fasta, size_t len, int k, int t, int nranks, uint64_t* stats, uint64_t* hist) {
  if (!valid_k(k) || nranks < 1 || nranks > 1024 || t < 1) return E_INVAL;
  RankRun rr;
  const int rc = k <= 32 ? pipeline_ranks<uint64_t>(fasta, len, k, t, nranks, rr)
                         : pipeline_ranks<u128>(fasta, len, k, t, nranks, rr);
  if (rc) return rc;
  stats[0] = rr.distinct;
  stats[1] = rr.kept;
  stats[2] = rr.nodes;
  stats[3] = rr.instances;
  stats[4] = (uint64_t)rr.min_bucket;
  for (int i = 0; i < t; i++) hist[i] = rr.hist[i];
  return rr.kept ? E_OK : E_EMPTY;
}

size_t pko_nodes_size(const pko_nodes* p) { return ((const Nodes*)p)->n_pre.size(); }
void pko_nodes_copy(const pko_nodes* p, uint64_t* keys, int32_t* pre_cnt, int32_t* suf_cnt,
                    uint8_t* pre_order, uint8_t* suf_order) {
  const Nodes* n = (const Nodes*)p;
  size_t m = n->n_pre.size();
  if (!m) return;
  if (keys) memcpy(keys, n->keys.data(), n->keys.size() * 8);
  if (pre_cnt) memcpy(pre_cnt, n->pre_cnt.data(), m * 16);
  if (suf_cnt) memcpy(suf_cnt, n->suf_cnt.data(), m * 16);
  if (pre_order) memcpy(pre_order, n->pre_order.data(), m * 4);
  if (suf_order) memcpy(suf_order, n->suf_order.data(), m * 4);
}
void pko_nodes_free(pko_nodes* p) { delete (Nodes*)p; }

}  // extern "C"
