// Drives pakhip::Hip_Kmer_Partial_Counter the way the reference's builder
// callers drive Concurrent_Kmer_Partial_Counter (src/default_kmer_table_builder.cpp:
// 16-81, src/concurrent_kmer_counter_test.cpp:82-114): an aggregating
// processor with atomic counters, one call of default_count_and_process.
//
//   builder_adapter_test                       reference unit-test input, klen 32
//   builder_adapter_test FILE KLEN READ_LEN BITS MAX_KMER_COUNT
//
//   builder_adapter_test v3 [FILE KLEN READ_LEN BITS MAX_KMER_COUNT]
//                                              the same through
//                                              concurrent_count_and_process_omp_v3,
//                                              called twice on one counter
//                                              (the reference unit test's call,
//                                              src/concurrent_kmer_counter_test.cpp:94)
//
//   builder_adapter_test explicit|implicit_v0|implicit_v1 FILE KLEN READ_LEN BITS
//                                              the counters src/concurrent_kmer_table_builder.cpp:
//                                              97-116 calls (processor never called):
//                                              their own "unique kmers=" and
//                                              "concurrent_vector size=" lines,
//                                              then "reads=R calls=C"
//
// Prints one line "reads=R unique=U freq=F calls=C xor=X" (X = xor of
// key words and counts over all pairs seen, an order-independent digest).
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "pakhip_builders.hpp"

struct count_unique_kmers_t {
  count_unique_kmers_t() : count_(0UL), kmer_freq_(0UL), calls_(0UL), digest_(0UL) {}
  template <typename iterator>
  void operator()(iterator beg, iterator end) const {
    ++calls_;
    for (; beg != end; ++beg) {
      ++count_;
      kmer_freq_ += beg->second;
      const unsigned __int128 w = (unsigned __int128)beg->first;
      digest_ ^= (unsigned long)(w ^ (w >> 64)) * 0x9E3779B97F4A7C15ull + (unsigned long)beg->second;
    }
  }
  mutable std::atomic<size_t> count_;
  mutable std::atomic<size_t> kmer_freq_;
  mutable std::atomic<size_t> calls_;
  mutable std::atomic<unsigned long> digest_;
};

template <class kmer_t>
int run(const std::string& input, size_t klen, size_t read_len, size_t max_kmer_count) {
  pakhip::Hip_Kmer_Partial_Counter<kmer_t> counter(klen, input.data(), input.data() + input.size(),
                                                   read_len, max_kmer_count);
  count_unique_kmers_t p;
  const size_t reads = counter.default_count_and_process(p);
  printf("reads=%zu unique=%zu freq=%zu calls=%zu xor=%lu\n", reads, (size_t)p.count_,
         (size_t)p.kmer_freq_, (size_t)p.calls_, (unsigned long)p.digest_);
  return 0;
}

// omp_v3 twice on one counter: it does not consume the range, so both calls
// see the same table; prints the second call's line.
template <class kmer_t>
int run_v3(const std::string& input, size_t klen, size_t read_len, size_t max_kmer_count) {
  pakhip::Hip_Kmer_Partial_Counter<kmer_t> counter(klen, input.data(), input.data() + input.size(),
                                                   read_len, max_kmer_count);
  count_unique_kmers_t p1, p2;
  const size_t r1 = counter.concurrent_count_and_process_omp_v3(p1);
  const size_t r2 = counter.concurrent_count_and_process_omp_v3(p2);
  if (r1 != r2 || (size_t)p1.count_ != (size_t)p2.count_ || (unsigned long)p1.digest_ != (unsigned long)p2.digest_) {
    printf("mismatch between calls\n");
    return 4;
  }
  printf("reads=%zu unique=%zu freq=%zu calls=%zu xor=%lu\n", r2, (size_t)p2.count_, (size_t)p2.kmer_freq_,
         (size_t)p2.calls_, (unsigned long)p2.digest_);
  return 0;
}

template <class kmer_t>
int run_table_builder_mode(const std::string& mode, const std::string& input, size_t klen, size_t read_len) {
  pakhip::Hip_Kmer_Partial_Counter<kmer_t> counter(klen, input.data(), input.data() + input.size(), read_len,
                                                   std::numeric_limits<size_t>::max());
  count_unique_kmers_t p;
  size_t reads;
  if (mode == "explicit") reads = counter.concurrent_count_and_process_omp_explicit(p);
  else if (mode == "implicit_v0") reads = counter.concurrent_count_and_process_omp_implicit_v0(p);
  else reads = counter.concurrent_count_and_process_omp_implicit_v1(p);
  printf("reads=%zu calls=%zu\n", reads, (size_t)p.calls_);
  return 0;
}

static std::string read_file(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) throw std::logic_error(std::string("Unable to open file: ") + path);
  std::string input;
  char buf[1 << 16];
  size_t r;
  while ((r = fread(buf, 1, sizeof buf, f)) > 0) input.append(buf, r);
  fclose(f);
  return input;
}

// pakgraph FILE K THRESHOLD COV: main-path count of FILE, then the PaK-graph
// builder nodes; one line per node:
//   key_lo key_hi | base:count:visit:term[:sidx/count/offset...] ... | suffixes
static int run_pakgraph(const char* path, int k, unsigned long threshold, int cov) {
  FILE* f = fopen(path, "rb");
  if (!f) throw std::logic_error(std::string("Unable to open file: ") + path);
  std::string input;
  char buf[1 << 16];
  size_t r;
  while ((r = fread(buf, 1, sizeof buf, f)) > 0) input.append(buf, r);
  fclose(f);
  pakhip::Device_Context ctx(0);
  pakhip::count_kmers(ctx.get(), input.data(), input.size(), k);
  auto nodes = pakhip::build_pakgraph_nodes(ctx.get(), threshold, cov);
  for (const auto& n : nodes) {
    printf("%lu %lu |", (unsigned long)n.key_lo, (unsigned long)n.key_hi);
    for (const auto& e : n.prefixes) {
      printf(" %c:%d:%d:%d", e.base ? e.base : '-', e.count, e.visit, (int)e.terminal);
      for (const auto& w : e.wires) printf(":%zu/%d/%d", w.suffix_idx, w.count, w.offset);
    }
    printf(" |");
    for (const auto& e : n.suffixes) printf(" %c:%d:%d:%d", e.base ? e.base : '-', e.count, e.visit, (int)e.terminal);
    printf("\n");
  }
  return 0;
}

int main(int argc, char** argv) {
  try {
    if (argc == 6 && std::string(argv[1]) == "pakgraph")
      return run_pakgraph(argv[2], atoi(argv[3]), strtoul(argv[4], nullptr, 10), atoi(argv[5]));
    // src/concurrent_kmer_counter_test.cpp:121-126
    const std::string unit_input =
        ">CACVBX020000001.1-NC_000913.3520\n"
        "TACCATTCACCCGCTGGCGGGGCAGGGGGTAAATCTCGGCTTTATGGATGCTGCAGAGCTGATTGCCGAACTGAAACGGTTGCATCGTCAGGGGAAAGAC\n"
        ">CACVBX020000001.1-NC_000913.3519\n"
        "TGGTTTGGCCGTCGCTTAGGCCAGGATGCCGCACCAGAGAAACTGGGCGTTGACCGCGCACTGGCATTCAGAGAAATCAATCAGCAACTTTATCAACTAC\n";
    if (argc == 6 && (std::string(argv[1]) == "explicit" || std::string(argv[1]) == "implicit_v0" ||
                      std::string(argv[1]) == "implicit_v1")) {
      const std::string input = read_file(argv[2]);
      const size_t klen = strtoull(argv[3], nullptr, 10), read_len = strtoull(argv[4], nullptr, 10);
      return atoi(argv[5]) == 64 ? run_table_builder_mode<size_t>(argv[1], input, klen, read_len)
                                 : run_table_builder_mode<unsigned __int128>(argv[1], input, klen, read_len);
    }
    if (argc >= 2 && std::string(argv[1]) == "v3") {
      if (argc == 2) return run_v3<size_t>(unit_input, 32, 100, std::numeric_limits<size_t>::max());
      if (argc != 7) {
        fprintf(stderr, "usage: %s v3 [FILE KLEN READ_LEN BITS MAX_KMER_COUNT]\n", argv[0]);
        return 2;
      }
      const std::string input = read_file(argv[2]);
      const size_t klen = strtoull(argv[3], nullptr, 10), read_len = strtoull(argv[4], nullptr, 10);
      const size_t maxc = strtoull(argv[6], nullptr, 10);
      return atoi(argv[5]) == 64 ? run_v3<size_t>(input, klen, read_len, maxc)
                                 : run_v3<unsigned __int128>(input, klen, read_len, maxc);
    }
    if (argc < 2) return run<size_t>(unit_input, 32, 100, std::numeric_limits<size_t>::max());
    if (argc < 6) {
      fprintf(stderr, "usage: %s FILE KLEN READ_LEN BITS MAX_KMER_COUNT\n", argv[0]);
      return 2;
    }
    const std::string input = read_file(argv[1]);
    const size_t klen = strtoull(argv[2], nullptr, 10), read_len = strtoull(argv[3], nullptr, 10);
    const int bits = atoi(argv[4]);
    const size_t maxc = strtoull(argv[5], nullptr, 10);
    return bits == 64 ? run<size_t>(input, klen, read_len, maxc)
                      : run<unsigned __int128>(input, klen, read_len, maxc);
  } catch (const std::logic_error& e) {
    printf("logic_error: %s\n", e.what());
    return 3;
  }
}
