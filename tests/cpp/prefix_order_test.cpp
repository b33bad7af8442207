// Host-only check of pakhip::reference_prefix_order and pakhip::visit_value
// (include/pakhip_builders.hpp).  Reads "k words n" then n kept keys (words
// u64 each, lo first) and m node keys from stdin (binary after the header
// line), prints the prefix bases of every node, then visit values.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pakhip_builders.hpp"

int main() {
  int k, words;
  unsigned long n, m;
  if (scanf("%d %d %lu %lu\n", &k, &words, &n, &m) != 4) return 2;
  pakhip::Kmer_Table kept;
  kept.k = k;
  kept.words = words;
  kept.keys.resize(n * words);
  kept.counts.assign(n, 1);
  if (fread(kept.keys.data(), 8, n * words, stdin) != n * words) return 3;
  std::vector<pk_mnode> nodes(m);
  for (unsigned long i = 0; i < m; i++) {
    uint64_t w[2] = {0, 0};
    if (fread(w, 8, words, stdin) != (size_t)words) return 4;
    nodes[i].key_lo = w[0];
    nodes[i].key_hi = words == 2 ? w[1] : 0;
  }
  auto order = pakhip::reference_prefix_order(kept, nodes);
  for (auto& o : order) printf("%d %d %d %d\n", o[0], o[1], o[2], o[3]);
  for (int c = 0; c < 300; c += 7)
    for (int cov : {1, 3, 30, 80}) printf("%d\n", pakhip::visit_value(c, cov));
  return 0;
}
