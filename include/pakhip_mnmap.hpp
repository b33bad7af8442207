// pakhip_mnmap.hpp -- fills the reference's macro-node map from the device's
// node records, for the reference's host-side consumers.
//
//   pakhip::fill_MN_map(nodes, n, k, coverage, pre_order, MN_map)
//       MN_map: std::vector<std::pair<kmer_t, MacroNode>> as
//       construct_macro_nodes leaves it (src/mn_construct.cpp:212-477, one
//       rank) and initiate_mnode_wiring (src/wire.cpp:66-121), compaction
//       and the walk take it.  Per node, in ascending key order:
//         k_1_mer     the key's k-1 bases, first base first (kmerel_mn,
//                     mn_construct.cpp:372-378)
//         suffixes    one single-base entry per present suffix, ascending
//                     base (the seq-sorted walk, mn_construct.cpp:344-391),
//                     terminal flag false, (count, visit_value(count))
//         prefixes    one single-base entry per present prefix in the order
//                     pre_order gives (pakhip::reference_prefix_order
//                     reproduces the reference's Comp_suffix order,
//                     mn_construct.cpp:299-303,395-457), or ascending base
//                     when pre_order is null; same flags and pairs
//       visit_value is the reference's ceil of a double quotient
//       (mn_construct.cpp:75-79).
//
// Header-only and templated on the consumer's types, so it compiles against
// the reference's own MacroNode / BasePairVector (includes/distribute_kmers.h:
// 114-395,533-658) without naming them: MacroNode needs the members the
// reference's has (k_1_mer, suffixes, suffixes_terminal, suffix_count and the
// prefix counterparts) and the base-pair vector a push_back of one 2-bit
// base.  Host code only.
#ifndef PAKHIP_MNMAP_HPP
#define PAKHIP_MNMAP_HPP

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

#include "pakhip.h"

namespace pakhip {

// visit_value (src/mn_construct.cpp:75-79).
inline int mn_visit_value(int count, int coverage) {
  return (int)std::ceil((double)count / (double)coverage);
}

// The node key of a record as the consumer's k-mer word (64- or 128-bit).
template <class kmer_t>
inline kmer_t mnode_key(const pk_mnode& r) {
  static_assert(sizeof(kmer_t) == 8 || sizeof(kmer_t) == 16, "kmer_t must be a 64- or 128-bit word");
  if (sizeof(kmer_t) == 8) return (kmer_t)r.key_lo;
  return (kmer_t)(((unsigned __int128)r.key_hi << 64) | (unsigned __int128)r.key_lo);
}

template <class kmer_t, class MacroNode>
void fill_MN_map(const pk_mnode* nodes, size_t n, int k, int coverage, const std::array<uint8_t, 4>* pre_order,
                 std::vector<std::pair<kmer_t, MacroNode>>& MN_map) {
  if (k < 2 || 2 * (k - 1) > (int)(8 * sizeof(kmer_t))) throw std::logic_error("fill_MN_map: k does not fit kmer_t");
  if (coverage < 1) throw std::logic_error("fill_MN_map: coverage must be >= 1");
  typedef decltype(MacroNode{}.k_1_mer) BasePairVector;
  const int L = k - 1;  // MN_LENGTH
  MN_map.clear();
  MN_map.reserve(n);
  for (size_t i = 0; i < n; i++) {
    const pk_mnode& r = nodes[i];
    const kmer_t key = mnode_key<kmer_t>(r);
    MN_map.emplace_back(key, MacroNode());
    MacroNode& mn = MN_map.back().second;
    BasePairVector k1;
    for (int p = 0; p < L; p++) k1.push_back((uint8_t)((key >> (2 * (L - p - 1))) & 3));
    mn.k_1_mer = k1;
    for (int b = 0; b < 4; b++) {
      if (!r.suf_cnt[b]) continue;
      BasePairVector s;
      s.push_back((uint8_t)b);
      mn.suffixes.push_back(s);
      mn.suffixes_terminal.push_back(false);
      mn.suffix_count.push_back(std::make_pair((int)r.suf_cnt[b], mn_visit_value(r.suf_cnt[b], coverage)));
    }
    for (int j = 0; j < 4; j++) {
      int a = j;
      if (pre_order) {
        a = pre_order[i][j];
        if (a == 255) break;
        if (a > 3 || !r.pre_cnt[a]) throw std::logic_error("fill_MN_map: prefix order names an absent prefix");
      } else if (!r.pre_cnt[a]) {
        continue;
      }
      BasePairVector p;
      p.push_back((uint8_t)a);
      mn.prefixes.push_back(p);
      mn.prefixes_terminal.push_back(false);
      mn.prefix_count.push_back(std::make_pair((int)r.pre_cnt[a], mn_visit_value(r.pre_cnt[a], coverage)));
    }
  }
}

}  // namespace pakhip

#endif  // PAKHIP_MNMAP_HPP
