// pk_gen.cpp -- seeded synthetic genomes and read sets written as strict
// single-line FASTA (">r<id>\n<sequence>\n"), the input format of PaKman's
// reader (src/kmer_counting.cpp:2920-2945 of the reference).
//
// Random numbers: splitmix64 seeds xoshiro256**.  Every read draws from its
// own stream (seeded from the read seed and the read index), so the output is
// identical for any thread count.
//
// Genome model: i.i.d. bases with P(C)=P(G)=gc/2, P(A)=P(T)=(1-gc)/2.
// Optionally a fraction rep_frac of the genome is overwritten by copies of
// rep_families random consensus sequences (length uniform in
// [rep_min, rep_max]), each copy carrying i.i.d. substitutions at rate rep_div.
// Reads: start uniform in [0, G-L], strand uniform (reverse complement on the
// minus strand), i.i.d. substitutions at rate err to one of the other three
// bases uniformly.  Read count R = round(G * coverage / L).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
};

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed) {
    SplitMix sm{seed};
    for (auto& x : s) x = sm.next();
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

const char BASES[4] = {'A', 'C', 'G', 'T'};

inline char comp(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return 'N';
  }
}

inline int base_index(char c) {
  switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    default: return 3;
  }
}

inline char draw_base(Xoshiro& r, double gc) {
  double u = r.uniform();
  if (u < gc) return (u < gc * 0.5) ? 'C' : 'G';
  return (u < gc + (1.0 - gc) * 0.5) ? 'A' : 'T';
}

inline char substitute(Xoshiro& r, char c) {
  int b = base_index(c);
  int o = (int)r.below(3);
  return BASES[(b + 1 + o) & 3];
}

inline size_t digits(uint64_t x) {
  size_t d = 1;
  while (x >= 10) { x /= 10; d++; }
  return d;
}

}  // namespace

extern "C" {

// Fill genome[0..G) (ASCII ACGT).
void pkgen_genome(uint64_t seed, uint64_t G, double gc, int rep_families, int rep_min,
                  int rep_max, double rep_frac, double rep_div, char* genome) {
  Xoshiro r(seed);
  for (uint64_t i = 0; i < G; i++) genome[i] = draw_base(r, gc);
  if (rep_families <= 0 || rep_frac <= 0.0 || G == 0) return;
  std::vector<std::vector<char>> fam((size_t)rep_families);
  for (auto& f : fam) {
    uint64_t len = (uint64_t)rep_min + r.below((uint64_t)(rep_max - rep_min + 1));
    if (len > G) len = G;
    f.resize(len);
    for (auto& c : f) c = draw_base(r, gc);
  }
  uint64_t target = (uint64_t)(rep_frac * (double)G);
  uint64_t placed = 0;
  while (placed < target) {
    const std::vector<char>& f = fam[r.below(fam.size())];
    uint64_t pos = r.below(G - f.size() + 1);
    for (size_t j = 0; j < f.size(); j++) {
      char c = f[j];
      if (r.uniform() < rep_div) c = substitute(r, c);
      genome[pos + j] = c;
    }
    placed += f.size();
  }
}

// Number of reads for a genome of G bases at the given coverage.
uint64_t pkgen_num_reads(uint64_t G, uint32_t L, double coverage) {
  if (L == 0 || G < L) return 0;
  return (uint64_t)llround((double)G * coverage / (double)L);
}

// Exact FASTA byte size of R reads of length L (">r<id>\n" + L + "\n").
uint64_t pkgen_fasta_size(uint64_t R, uint32_t L) {
  uint64_t total = 0;
  uint64_t lo = 0, p10 = 10;
  size_t d = 1;
  while (lo < R) {
    uint64_t hi = p10 < R ? p10 : R;
    total += (hi - lo) * (3 + d + (uint64_t)L + 1);
    lo = hi;
    p10 *= 10;
    d++;
  }
  return total;
}

// Write R reads sampled from genome[0..G) into out (pkgen_fasta_size bytes).
void pkgen_reads(uint64_t seed, const char* genome, uint64_t G, uint32_t L, uint64_t R,
                 double err, char* out) {
  const uint64_t span = G - L + 1;
#pragma omp parallel for schedule(static, 4096)
  for (long long i = 0; i < (long long)R; i++) {
    uint64_t d = digits((uint64_t)i);
    // offset of record i: all records before it (header digits differ).
    uint64_t off = pkgen_fasta_size((uint64_t)i, L);
    char* w = out + off;
    *w++ = '>';
    *w++ = 'r';
    char num[24];
    snprintf(num, sizeof num, "%lld", i);
    memcpy(w, num, d);
    w += d;
    *w++ = '\n';
    Xoshiro r(seed ^ (0x9e3779b97f4a7c15ull * (uint64_t)(i + 1)));
    uint64_t pos = r.below(span);
    bool minus = (r.next() >> 63) != 0;
    for (uint32_t j = 0; j < L; j++) {
      char c = minus ? comp(genome[pos + L - 1 - j]) : genome[pos + j];
      if (err > 0.0 && r.uniform() < err) c = substitute(r, c);
      w[j] = c;
    }
    w[L] = '\n';
  }
}

}  // extern "C"
