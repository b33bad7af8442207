// ref_harness.cpp -- drives the reference's own histogram cutoff, macro-node
// construction, wiring, compaction and walk (compiled from the reference
// tree by oracle/build_ref.sh) on a brute-force k-mer table.
//
// TEST INFRASTRUCTURE ONLY (builds oracle/_ref/pakref_k<K>).  The reference's
// counting TU (src/kmer_counting.cpp) cannot be built in this image: it
// includes oneTBB headers (src/concurrent_kmer_counter.hpp:11-14) that are not
// installed.  This harness therefore supplies the two symbols that TU would
// define for the rest of the pipeline:
//   * perform_kmer_counting(): fills the global kmer_proc_buf with the
//     sorted, unique (k-mer, count) table, counted here by brute force
//     (std::map over every window of every read line, read with the record
//     rules of src/kmer_counting.cpp:2920-2945);
//   * free_kmer_count_buffers(): clears kmer_proc_buf
//     (src/kmer_counting.cpp:3343-3349).
// Everything after counting is the reference's own code, called in the order
// of src/distribute_kmers.cpp:132-215.  Extra dumps (macro nodes before and
// after wiring) go to files named by $PAKREF_DUMP (default "pakref").
//
// With $PAKREF_NODES naming a file of node records (written by the tests:
// "PKMN", u64 count, i32 k, then count pk_mnode records and count x 4
// prefix-order bytes), counting and begin_mnode_construction are skipped:
// the MN_map is filled from those records by the library's filler
// (include/pakhip_mnmap.hpp, pakhip::fill_MN_map) and the reference's
// wiring, compaction and walk run on it.  The contigs then show whether the
// records plus the filler are a drop-in for construct_macro_nodes.  The
// fill's time is printed; $PAKREF_FILL_ONLY stops after the fill.

#include <array>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "distribute_kmers.h"
#include "pakhip_mnmap.hpp"

extern std::vector<KmerPairs> kmer_proc_buf;
extern std::string inputFileName;
extern int rank, size;
extern int coverage;
void parseCommandLine(const int argc, char* const argv[]);
void set_num_threads();

void free_kmer_count_buffers() {
  kmer_proc_buf.clear();
  kmer_proc_buf.shrink_to_fit();
}

void perform_kmer_counting(const char* s, size_t n) {
  const size_t k = KMER_LENGTH;
  std::map<kmer_t, long long> m;
  size_t i = 0;
  while (i < n && s[i] != '>') i++;
  while (i < n) {
    const char* nl = (const char*)memchr(s + i, '\n', n - i);
    if (!nl) break;
    size_t b = (size_t)(nl - s) + 1;
    if (b + k > n) break;
    const char* nl2 = (const char*)memchr(s + b, '\n', n - b);
    size_t end = nl2 ? (size_t)(nl2 - s) : n;
    for (size_t w = b; w + k <= end; w++) {
      kmer_t v = 0;
      for (size_t j = 0; j < k; j++) v = kmer_shift(v, char_to_el(s[w + j]));
      m[v]++;
    }
    i = end + 1;
  }
  kmer_proc_buf.clear();
  for (auto& kv : m) kmer_proc_buf.push_back(KmerPairs{kv.first, (int)kv.second});
  if (rank == 0) printf("Total distinct k-mer entries across all proc's: %lu\n", (unsigned long)m.size());
}

static void put_key(FILE* f, kmer_t x) {
  unsigned long long lo = (unsigned long long)x;
  unsigned long long hi = 0;
#ifdef EXTEND_KMER
  hi = (unsigned long long)(x >> 64);
#endif
  fprintf(f, "%016llx %016llx", lo, hi);
}

static void dump_nodes(const std::vector<std::pair<kmer_t, MacroNode>>& mn, const std::string& path,
                       bool wiring) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) { perror(path.c_str()); exit(3); }
  for (const auto& e : mn) {
    const MacroNode& nd = e.second;
    fprintf(f, "N ");
    put_key(f, e.first);
    std::string k1;
    for (size_t l = 0; l < nd.k_1_mer.size(); l++) k1 += el_to_char(nd.k_1_mer[l]);
    fprintf(f, " %s P %zu", k1.c_str(), nd.prefixes.size());
    for (size_t l = 0; l < nd.prefixes.size(); l++) {
      std::string x;
      for (size_t t = 0; t < nd.prefixes[l].size(); t++) x += el_to_char(nd.prefixes[l][t]);
      fprintf(f, " %s %d %d %d", x.empty() ? "-" : x.c_str(), nd.prefix_count[l].first,
              nd.prefix_count[l].second, (int)nd.prefixes_terminal[l]);
    }
    fprintf(f, " S %zu", nd.suffixes.size());
    for (size_t l = 0; l < nd.suffixes.size(); l++) {
      std::string x;
      for (size_t t = 0; t < nd.suffixes[l].size(); t++) x += el_to_char(nd.suffixes[l][t]);
      fprintf(f, " %s %d %d %d", x.empty() ? "-" : x.c_str(), nd.suffix_count[l].first,
              nd.suffix_count[l].second, (int)nd.suffixes_terminal[l]);
    }
    if (wiring) {
      fprintf(f, " W %zu", nd.wiring_info.size());
      for (const auto& w : nd.wiring_info)
        fprintf(f, " %d %d %d", w.suffix_id, w.offset_in_suffix, w.count);
      fprintf(f, " B %zu", nd.prefix_begin_info.size());
      for (const auto& b : nd.prefix_begin_info) fprintf(f, " %d %d", b.prefix_pos, b.num_wires);
    }
    fprintf(f, "\n");
  }
  fclose(f);
}

// MN_map from a node-record file (see the header comment); false on a bad file.
static bool fill_from_records(const char* path, std::vector<std::pair<kmer_t, MacroNode>>& MN_map) {
  FILE* f = fopen(path, "rb");
  if (!f) { perror(path); return false; }
  char magic[4];
  unsigned long long n = 0;
  int k = 0;
  bool ok = fread(magic, 1, 4, f) == 4 && memcmp(magic, "PKMN", 4) == 0 && fread(&n, 8, 1, f) == 1 &&
            fread(&k, 4, 1, f) == 1 && k == KMER_LENGTH;
  std::vector<pk_mnode> recs(ok ? n : 0);
  std::vector<std::array<uint8_t, 4>> order(ok ? n : 0);
  if (ok && n) ok = fread(recs.data(), sizeof(pk_mnode), n, f) == n && fread(order.data(), 4, n, f) == n;
  fclose(f);
  if (!ok) { fprintf(stderr, "%s: not a node-record file for k=%d\n", path, KMER_LENGTH); return false; }
  const double t0 = MPI_Wtime();
  pakhip::fill_MN_map<kmer_t, MacroNode>(recs.data(), recs.size(), k, coverage, order.data(), MN_map);
  const double t1 = MPI_Wtime();
  if (rank == 0) {
    printf("Total number of macro_nodes across all proc's: %lu\n", (unsigned long)MN_map.size());
    printf("MN_map fill from node records (secs): %f\n", t1 - t0);
  }
  return true;
}

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &size);
  parseCommandLine(argc, argv);
  set_num_threads();
  const char* dump = getenv("PAKREF_DUMP");
  std::string prefix = dump ? dump : "pakref";

  std::vector<std::pair<kmer_t, MacroNode>> MN_map;
  const char* recs = getenv("PAKREF_NODES");
  if (recs) {
    if (!fill_from_records(recs, MN_map)) return 2;
    if (getenv("PAKREF_FILL_ONLY")) {
      MPI_Finalize();
      return 0;
    }
  } else {
  FILE* fp = fopen(inputFileName.c_str(), "rb");
  if (!fp) { perror(inputFileName.c_str()); return 2; }
  fseek(fp, 0, SEEK_END);
  long sz = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  std::vector<char> buf((size_t)sz + 1, '\0');
  if (fread(buf.data(), 1, (size_t)sz, fp) != (size_t)sz) { fclose(fp); return 2; }
  fclose(fp);

  perform_kmer_counting(buf.data(), (size_t)sz);

  begin_mnode_construction(MN_map);
  }
  dump_nodes(MN_map, prefix + "_mn.txt", false);
  initiate_mnode_wiring(MN_map);
  dump_nodes(MN_map, prefix + "_wired.txt", true);

  if (!getenv("PAKREF_NO_WALK")) {
    std::vector<BasePairVector> partial_contig_list;
    size_t global_num_nodes = begin_iterative_compaction(MN_map, partial_contig_list);
    MPI_Barrier(MPI_COMM_WORLD);
    std::vector<BeginMN> list_of_begin_kmers;
    identify_begin_kmers(MN_map, list_of_begin_kmers);
    std::vector<std::pair<kmer_t, MacroNode>> global_MN_map(global_num_nodes);
    generate_compacted_pakgraph(MN_map, global_MN_map);
    traverse_pakgraph(global_MN_map, list_of_begin_kmers, partial_contig_list);
  }
  MPI_Barrier(MPI_COMM_WORLD);
  MPI_Finalize();
  return 0;
}
