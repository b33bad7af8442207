// pk_device.hip -- MI355X (gfx950) k-mer counting, histogram cutoff and
// macro-node construction behind the C ABI of include/pakhip.h.
//
// Pipeline of pk_count (one HIP stream per context, everything resident in HBM):
//   1. line index     k_nl_count / scan / k_nl_write: positions of every '\n'
//                     (16 B per thread, block-scan compaction).
//   2. read table     k_reads: record r = header line j0+2r, sequence line
//                     j0+2r+1, with the record rules and the format check of
//                     the reference's Sliding_window (src/kmer_counting.cpp:
//                     2890-2945); k-mer offsets by exclusive scan.  The
//                     builder API's records come from k_bld_flags/k_bld_write.
//   3. partition      multisplit by the top bits of the k-mer.  Level 1
//                     (k_l1_hist + k_l1_scatter, <= 10 bits) is fused with the
//                     parse: each chunk of reads is staged in LDS and every
//                     read is walked by one thread with the reference's rolling
//                     3-bit update (kmer_shift, distribute_kmers.h:798-805);
//                     every k-mer is stored at its block's next slot for its
//                     digit.  Further levels (k_ms_hist + k_ms_scatter, <= 8
//                     bits) split each bucket tile by tile, sorted by digit in
//                     LDS and written as contiguous runs.  This is the on-device
//                     stand-in for the reference's owner routing +
//                     MPI_Alltoallv (kmer_counting.cpp:254-352); rank-share
//                     mode keeps one owner's k-mers (own_step).
//   4. bucket count   k_bucket_hash: one 256-thread workgroup per bucket counts
//                     the bucket's distinct keys in an LDS hash table, then
//                     counting-sorts them by the next bits and writes the
//                     ascending (key, count) pairs (the reference's sort + RLE
//                     + merge + aggregate, kmer_counting.cpp:3057-3086,
//                     169-250).  Buckets with more distinct keys than a table
//                     holds take extra passes, then the host-driven split
//                     k_part_hist / k_part_scatter.
//   5. compaction     k_compact: bucket results -> dense sorted table.
// pk_cutoff: per-block LDS histograms of counts < t, k_min_bucket, ordered
//   stream compaction (kmer_histogram.cpp:84-140).
// pk_macro_nodes: range-partitioned 5-way merge of the kept table seen as
//   pred keys (ascending) and, per first base a, succ keys (ascending within
//   the quarter) -- no sort is needed (mn_construct.cpp:212-477).

#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pakhip.h"

typedef unsigned __int128 u128;
typedef unsigned long long ull;

// ---------------------------------------------------------------------------
// Key helpers (Key = uint64_t for k <= 32, u128 for 32 < k <= 64).

template <class Key>
__host__ __device__ __forceinline__ Key key_mask(int nbits) {
  const int bits = (int)sizeof(Key) * 8;
  if (nbits >= bits) return ~Key(0);
  return (Key(1) << nbits) - Key(1);
}

// bits [pos, pos+nbits) of x; bits below 0 read as 0.  nbits <= 16.
template <class Key>
__device__ __forceinline__ uint32_t key_digit(Key x, int pos, int nbits) {
  if (nbits <= 0) return 0;
  Key y;
  if (pos >= 0) {
    if (pos >= (int)sizeof(Key) * 8) return 0;
    y = x >> pos;
  } else {
    y = x << (-pos);
  }
  return (uint32_t)y & ((1u << nbits) - 1u);
}

// u128 keys: the digit from a 32-bit funnel shift of the two 32-bit words
// around pos (v_alignbit), the words picked by a uniform pos without
// branches; same value as the generic form for nbits <= 32.
template <>
__device__ __forceinline__ uint32_t key_digit<unsigned __int128>(unsigned __int128 x, int pos, int nbits) {
  if (nbits <= 0) return 0;
  if (pos < 0) return (uint32_t)(x << (-pos)) & ((1u << nbits) - 1u);
  if (pos >= 128) return 0;
  const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  const uint32_t w0 = (uint32_t)lo, w1 = (uint32_t)(lo >> 32), w2 = (uint32_t)hi, w3 = (uint32_t)(hi >> 32);
  const int wi = pos >> 5;
  const uint32_t a = wi == 0 ? w0 : wi == 1 ? w1 : wi == 2 ? w2 : w3;
  const uint32_t b = wi == 0 ? w1 : wi == 1 ? w2 : wi == 2 ? w3 : 0u;
  return __builtin_amdgcn_alignbit(b, a, (uint32_t)(pos & 31)) & ((1u << nbits) - 1u);
}

// ---------------------------------------------------------------------------
// Block-wide primitives (wave64).

// Exclusive prefix sum of v over the block (all NT threads must call it).
template <int NT>
__device__ __forceinline__ ull block_excl_scan(ull v) {
  constexpr int NW = NT / 64;
  __shared__ ull wsum[NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  ull x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    ull y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    const ull s = lane < NW ? wsum[lane] : 0ull;
    ull inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      ull y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    if (lane < NW) wsum[lane] = inc - s;
  }
  __syncthreads();
  const ull r = x - v + wsum[w];
  __syncthreads();
  return r;
}

// Exclusive prefix sum of v over the block and the block total, without a
// trailing barrier: the caller must pass a barrier before the next call.
template <int NT>
__device__ __forceinline__ ull block_scan_total(ull v, ull* total) {
  constexpr int NW = NT / 64;
  __shared__ ull wsum[NW + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  ull x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    ull y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    const ull s = lane < NW ? wsum[lane] : 0ull;
    ull inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      ull y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    if (lane < NW) wsum[lane] = inc - s;
    if (lane == NW - 1) wsum[NW] = inc;
  }
  __syncthreads();
  *total = wsum[NW];
  return x - v + wsum[w];
}

template <int NT>
__device__ __forceinline__ ull block_sum(ull v) {
  constexpr int NW = NT / 64;
  __shared__ ull ws[NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) ws[w] = v;
  __syncthreads();
  ull s = 0;
#pragma unroll
  for (int i = 0; i < NW; i++) s += ws[i];
  __syncthreads();
  return s;
}

template <int NT>
__device__ __forceinline__ ull block_min(ull v) {
  constexpr int NW = NT / 64;
  __shared__ ull ws[NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ull y = __shfl_xor(v, o, 64);
    v = y < v ? y : v;
  }
  if (lane == 0) ws[w] = v;
  __syncthreads();
  ull s = ~0ull;
#pragma unroll
  for (int i = 0; i < NW; i++) s = ws[i] < s ? ws[i] : s;
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// Small utility kernels.

__global__ void k_fill_u64(ull* p, ull v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_copy_u64(ull* dst, const ull* src) { *dst = *src; }

// out[i] = off[i + 1] - off[i]
__global__ void k_diff(const ull* off, uint64_t n, ull* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = off[i + 1] - off[i];
}

// ---------------------------------------------------------------------------
// Exclusive scan (u32 or u64 in, u64 out, out[n] = total).  SCAN_PER
// elements per 1024-thread block, SCAN_EPT consecutive ones per thread
// (read and written as 16-byte vectors where aligned, so a thread's loads
// are in flight together), recursive over block sums.

constexpr int SCAN_NT = 1024;
constexpr int SCAN_EPT = 16;
constexpr int SCAN_PER = SCAN_NT * SCAN_EPT;

// elements [base, base + SCAN_EPT) of in, zero past n
template <class T>
__device__ __forceinline__ void scan_load(const T* in, uint64_t n, uint64_t base, ull* v) {
  if (base + SCAN_EPT <= n && (reinterpret_cast<uintptr_t>(in + base) & 15) == 0) {
    constexpr int PV = 16 / (int)sizeof(T);  // elements per 16-byte vector
    const uint4* p = reinterpret_cast<const uint4*>(in + base);
    uint4 q[SCAN_EPT / PV];
#pragma unroll
    for (int j = 0; j < SCAN_EPT / PV; j++) q[j] = p[j];
#pragma unroll
    for (int j = 0; j < SCAN_EPT / PV; j++) {
      const T* e = reinterpret_cast<const T*>(&q[j]);
#pragma unroll
      for (int i = 0; i < PV; i++) v[j * PV + i] = (ull)e[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_EPT; i++) v[i] = base + i < n ? (ull)in[base + i] : 0ull;
  }
}

template <class T>
__global__ __launch_bounds__(SCAN_NT) void k_scan_reduce(const T* in, ull* sums, uint64_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_PER + (uint64_t)threadIdx.x * SCAN_EPT;
  ull v[SCAN_EPT];
  scan_load<T>(in, n, base, v);
  ull s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_EPT; i++) s += v[i];
  s = block_sum<SCAN_NT>(s);
  if (threadIdx.x == 0) sums[blockIdx.x] = s;
}

template <class T>
__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(const T* in, ull* out, const ull* offs,
                                                        uint64_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_PER + (uint64_t)threadIdx.x * SCAN_EPT;
  ull v[SCAN_EPT];
  scan_load<T>(in, n, base, v);
  ull s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_EPT; i++) s += v[i];
  const ull ex = block_excl_scan<SCAN_NT>(s);
  ull o = (offs ? offs[blockIdx.x] : 0ull) + ex;
  if (base + SCAN_EPT <= n && (reinterpret_cast<uintptr_t>(out + base) & 15) == 0) {
    ulonglong2* q = reinterpret_cast<ulonglong2*>(out + base);
#pragma unroll
    for (int j = 0; j < SCAN_EPT / 2; j++) {
      const ull a = o;
      o += v[2 * j];
      q[j] = make_ulonglong2(a, o);
      o += v[2 * j + 1];
    }
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_EPT; i++) {
      if (base + i < n) out[base + i] = o;
      o += v[i];
    }
  }
  // the thread owning element n-1 writes the total
  if (n > 0 && base <= n - 1 && n - 1 < base + SCAN_EPT) out[n] = o;
}

// ---------------------------------------------------------------------------
// Line index.

constexpr int NL_NT = 256;
// A block covers NL_SLABS slabs of NL_SLAB bytes; in slab j lane t reads the
// 16 bytes at 16 t (one coalesced 16-byte load per lane and slab).
constexpr int NL_SLABS = 4;
constexpr uint64_t NL_SLAB = (uint64_t)NL_NT * 16;
constexpr uint64_t NL_PER_BLOCK = NL_SLAB * NL_SLABS;

// 4-bit mask of the bytes of w equal to e (bit i = byte i): a byte is zero
// in t = w ^ eeee exactly when bit 7 of ((t & 7f..) + 7f..) | t is clear.
__device__ __forceinline__ uint32_t match4(uint32_t w, uint32_t e) {
  const uint32_t t = w ^ (e * 0x01010101u);
  const uint32_t z = ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u;
  return ((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u);
}

// Masks (bit i = byte base + i) of the event bytes (e0 or e1) and, when
// want_gt, of the '>' bytes among the 16 bytes at base (bytes at or past n
// never match); returns the 16 bytes (zero past n).  The kernels are bound
// by these byte matches, so a second match runs only when e0 != e1 (the
// builder's delimiter and skip symbol; the main path looks for '\n' alone)
// and the '>' match only where a caller needs it.
__device__ __forceinline__ uint4 scan16(const uint8_t* fa, uint64_t n, uint64_t base, uint32_t e0, uint32_t e1,
                                        bool want_gt, uint32_t& ev, uint32_t& gt) {
  ev = 0;
  gt = 0;
  if (base + 16 <= n && ((reinterpret_cast<uintptr_t>(fa + base) & 15) == 0)) {
    const uint4 v = *reinterpret_cast<const uint4*>(fa + base);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      uint32_t m = match4(w[i], e0);
      if (e1 != e0) m |= match4(w[i], e1);
      ev |= m << (4 * i);
      if (want_gt) gt |= match4(w[i], '>') << (4 * i);
    }
    return v;
  }
  uint32_t w[4] = {0, 0, 0, 0};
  for (int i = 0; i < 16; i++) {
    if (base + i >= n) break;
    const uint32_t b = fa[base + i];
    w[i >> 2] |= b << (8 * (i & 3));
    ev |= (uint32_t)(b == e0 || b == e1) << i;
    gt |= (uint32_t)(b == '>') << i;
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// scan16 that also returns the 16 bytes (zero past n).
__device__ __forceinline__ uint4 scan16_keep(const uint8_t* in, uint64_t n, uint64_t base, uint32_t e0,
                                             uint32_t& ev) {
  ev = 0;
  uint4 v = make_uint4(0, 0, 0, 0);
  if (base + 16 <= n && ((reinterpret_cast<uintptr_t>(in + base) & 15) == 0)) {
    v = *reinterpret_cast<const uint4*>(in + base);
    ev = match4(v.x, e0) | (match4(v.y, e0) << 4) | (match4(v.z, e0) << 8) | (match4(v.w, e0) << 12);
  } else {
    uint32_t w[4] = {0, 0, 0, 0};
    for (int i = 0; i < 16; i++) {
      if (base + i >= n) break;
      const uint32_t b = in[base + i];
      w[i >> 2] |= b << (8 * (i & 3));
      ev |= (uint32_t)(b == e0) << i;
    }
    v = make_uint4(w[0], w[1], w[2], w[3]);
  }
  return v;
}

__device__ __forceinline__ uint8_t byte_of(const uint4& v, int i) {
  const uint32_t w = i < 8 ? (i < 4 ? v.x : v.y) : (i < 12 ? v.z : v.w);
  return (uint8_t)(w >> (8 * (i & 3)));
}

// Event bytes of the index: bytes equal to e0 or e1 ('\n' and '\n' for the
// line index of the main path; delimiter and skip symbol for the builder).
// Per block: the number of event bytes and the first '>'.  Only the input's
// first '>' is wanted, so only block 0 looks for it unless gt_all is set
// (the host asks for that pass when block 0 holds no '>').
__global__ __launch_bounds__(NL_NT) void k_nl_count(const uint8_t* fa, uint64_t n, uint8_t e0,
                                                    uint8_t e1, ull* blk_cnt, ull* blk_gt, int gt_all) {
  const uint64_t b0 = (uint64_t)blockIdx.x * NL_PER_BLOCK + (uint64_t)threadIdx.x * 16;
  const bool want_gt = gt_all || blockIdx.x == 0;
  uint32_t ev[NL_SLABS], gm[NL_SLABS];
#pragma unroll
  for (int j = 0; j < NL_SLABS; j++) scan16(fa, n, b0 + j * NL_SLAB, e0, e1, want_gt, ev[j], gm[j]);
  ull c = 0, gt = ~0ull;
#pragma unroll
  for (int j = NL_SLABS - 1; j >= 0; j--) {
    c += __popc(ev[j]);
    if (gm[j]) gt = b0 + j * NL_SLAB + __ffs(gm[j]) - 1;
  }
  c = block_sum<NL_NT>(c);
  gt = block_min<NL_NT>(gt);
  if (threadIdx.x == 0) {
    blk_cnt[blockIdx.x] = c;
    blk_gt[blockIdx.x] = gt;
  }
}

// out[0] = min(in[0..n)) (one block).
__global__ __launch_bounds__(1024) void k_min_reduce(const ull* in, uint64_t n, ull* out) {
  ull m = ~0ull;
  for (uint64_t i = threadIdx.x; i < n; i += 1024) m = in[i] < m ? in[i] : m;
  m = block_min<1024>(m);
  if (threadIdx.x == 0) *out = m;
}

// Positions of the event bytes, ascending, from the block's base blk_off.
// The four slab counts of a thread travel through one block scan packed in
// 16-bit fields (a slab holds at most 4096 events, so no field carries).
// With follow = 1, bits 62-63 of each entry tell what follows the event: 1
// a '>', 2 the end of the input, 0 anything else (the record check would
// otherwise read that byte per record, one scattered line each).
constexpr int NL_FLAG_SHIFT = 62;
constexpr ull NL_POS_MASK = (1ull << NL_FLAG_SHIFT) - 1;

__global__ __launch_bounds__(NL_NT) void k_nl_write(const uint8_t* fa, uint64_t n, uint8_t e0,
                                                    uint8_t e1, const ull* blk_off, ull* nl, int follow) {
  const uint64_t b0 = (uint64_t)blockIdx.x * NL_PER_BLOCK + (uint64_t)threadIdx.x * 16;
  uint32_t ev[NL_SLABS];
  uint4 by[NL_SLABS];
  ull packed = 0;
#pragma unroll
  for (int j = 0; j < NL_SLABS; j++) {
    uint32_t gm;
    by[j] = scan16(fa, n, b0 + j * NL_SLAB, e0, e1, false, ev[j], gm);
    packed |= (ull)__popc(ev[j]) << (16 * j);
  }
  ull tot;
  const ull ex = block_scan_total<NL_NT>(packed, &tot);
  ull slab_base = blk_off[blockIdx.x];
#pragma unroll
  for (int j = 0; j < NL_SLABS; j++) {
    ull o = slab_base + ((ex >> (16 * j)) & 0xFFFFull);
    const uint64_t pos = b0 + j * NL_SLAB;
    for (uint32_t m = ev[j]; m; m &= m - 1) {
      const int i = __ffs(m) - 1;
      const uint64_t p = pos + i;
      ull f = 0;
      if (follow) {
        // the next byte: in this thread's 16 bytes, else the next chunk's
        // first (the neighbour's load, usually still in cache)
        f = p + 1 >= n ? 2ull : i < 15 ? (ull)(byte_of(by[j], i + 1) == '>') : (ull)(fa[p + 1] == '>');
      }
      nl[o++] = p | (f << NL_FLAG_SHIFT);
    }
    slab_base += (tot >> (16 * j)) & 0xFFFFull;
  }
}

// ---------------------------------------------------------------------------
// Input formats: multi-line FASTA and four-line FASTQ (either with '\r\n'
// line ends) rewritten into the strict single-line FASTA the counting path
// reads.  Line j of the input is [a, b): a = 0 or one past the (j-1)-th
// '\n', b = the j-th '\n' or the end; a '\r' at b - 1 is dropped.  One
// thread per line computes the line's output length, a scan places the
// lines, and a byte-parallel pass (k_fmt_copy) moves the bytes.

constexpr int FMT_FASTA = 1;
constexpr int FMT_FASTQ = 2;

__device__ __forceinline__ void fmt_line(const uint8_t* in, uint64_t n, const ull* nl, uint64_t m, uint64_t j,
                                         ull& a, ull& b) {
  a = j ? nl[j - 1] + 1 : 0;
  b = j < m ? nl[j] : n;
  if (b > a && in[b - 1] == '\r') b--;
}

// Multi-line FASTA: *hmin = the first line that starts with '>'.
__global__ void k_fmt_first_record(const uint8_t* in, uint64_t n, const ull* nl, uint64_t m, uint64_t L,
                                   ull* hmin) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  const ull a = j ? nl[j - 1] + 1 : 0;
  if (a < n && in[a] == '>') atomicMin(hmin, (ull)j);
}

// Output bytes of every line.  FASTA: nothing before the first record; a
// header line gives itself + '\n'; a sequence line gives itself; the last
// line of a record adds the '\n' that ends the read (a header with no
// sequence line gives an empty read).  FASTQ: header (its '@' becomes '>')
// and sequence lines give themselves + '\n'; the '+' and quality lines give
// nothing.  bad receives the first malformed FASTQ line, too_long the first
// line whose output exceeds 32 bits.
template <int FMT>
__global__ void k_fmt_lines(const uint8_t* in, uint64_t n, const ull* nl, uint64_t m, uint64_t L,
                            const ull* hmin, uint32_t* olen, ull* lstart, uint32_t* llen, ull* bad,
                            ull* too_long) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  ull a, b;
  fmt_line(in, n, nl, m, j, a, b);
  const ull len = b - a;
  lstart[j] = a;
  llen[j] = (uint32_t)(len < 0xFFFFFFFFull ? len : 0xFFFFFFFFull);
  ull out;
  if (FMT == FMT_FASTA) {
    if (j < *hmin) {
      out = 0;
    } else {
      const bool header = len > 0 && in[a] == '>';
      bool last = j + 1 >= L;
      if (!last) {
        const ull a2 = nl[j] + 1;
        last = a2 < n && in[a2] == '>';
      }
      out = len + (header ? 1 : 0) + (last ? 1 : 0);
    }
  } else {
    const int role = (int)(j & 3);
    if ((role == 0 && !(len > 0 && in[a] == '@')) || (role == 2 && !(len > 0 && in[a] == '+')))
      atomicMin(bad, (ull)j);
    out = role < 2 ? len + 1 : 0;
  }
  if (out > 0xFFFFFFFFull) {
    atomicMin(too_long, (ull)j);
    out = 0;
  }
  olen[j] = (uint32_t)out;
}

// The copy, byte-parallel in the layout of k_nl_write: lane t of slab j
// holds 16 input bytes, and the block scan of the '\n' counts gives the
// line of its first byte (blk_nl = '\n's before each block).  A content
// byte of line i goes to ooff[i] + (its offset in the line); the '\n' that
// ends line i (or the input's last byte, for a last line without one) writes
// the '\n's the line's output length adds after its content.  Sixteen
// content bytes of one line (most chunks) take one 16-byte store.
template <int FMT>
__global__ __launch_bounds__(NL_NT) void k_fmt_copy(const uint8_t* __restrict__ in, uint64_t n,
                                                    const ull* __restrict__ blk_nl,
                                                    const ull* __restrict__ lstart,
                                                    const uint32_t* __restrict__ llen,
                                                    const uint32_t* __restrict__ olen,
                                                    const ull* __restrict__ ooff, uint8_t* __restrict__ out) {
  const uint64_t b0 = (uint64_t)blockIdx.x * NL_PER_BLOCK + (uint64_t)threadIdx.x * 16;
  uint32_t ev[NL_SLABS];
  uint4 raw[NL_SLABS];
  ull packed = 0;
#pragma unroll
  for (int j = 0; j < NL_SLABS; j++) {
    raw[j] = scan16_keep(in, n, b0 + j * NL_SLAB, '\n', ev[j]);
    packed |= (ull)__popc(ev[j]) << (16 * j);
  }
  ull tot;
  const ull ex = block_scan_total<NL_NT>(packed, &tot);
  const ull blk_base = blk_nl[blockIdx.x];
  // slab: 16 bytes v at pos0 whose '\n's are evj, the first in line `line`
  auto slab = [&](uint64_t pos0, uint4 v, uint32_t evj, uint64_t line) {
    if (pos0 >= n) return;
    const int cnt = n - pos0 < 16 ? (int)(n - pos0) : 16;
    // the line's fields: start, content length, output length and place
    ull a = lstart[line], oo = ooff[line];
    uint32_t len = llen[line], ol = olen[line];
    if (cnt == 16 && evj == 0 && pos0 + 16 <= a + len && pos0 != n - 16) {
      // (a byte-aligned 16-byte copy: one unaligned dwordx4 store)
      if (ol && !(FMT == FMT_FASTQ && (line & 3) == 0 && pos0 == a)) __builtin_memcpy(out + oo + (pos0 - a), &v, 16);
      return;
    }
    for (int i = 0; i < cnt; i++) {
      const uint64_t pos = pos0 + i;
      const bool is_nl = (evj >> i) & 1u;
      if (ol) {
        if (pos < a + len) {
          const bool gt = FMT == FMT_FASTQ && (line & 3) == 0 && pos == a;
          out[oo + (pos - a)] = gt ? (uint8_t)'>' : byte_of(v, i);
        }
        if (is_nl || pos == n - 1) {
          // the end of the line: the '\n's after its content
          for (uint32_t q = len; q < ol; q++) out[oo + q] = '\n';
        }
      }
      if (is_nl && pos + 1 < n) {
        line++;
        a = lstart[line];
        oo = ooff[line];
        len = llen[line];
        ol = olen[line];
      }
    }
  };
  static_assert(NL_SLABS == 4, "four slabs");
  const ull t0 = tot & 0xFFFFull, t1 = (tot >> 16) & 0xFFFFull, t2 = (tot >> 32) & 0xFFFFull;
  slab(b0, raw[0], ev[0], blk_base + (ex & 0xFFFFull));
  slab(b0 + NL_SLAB, raw[1], ev[1], blk_base + t0 + ((ex >> 16) & 0xFFFFull));
  slab(b0 + 2 * NL_SLAB, raw[2], ev[2], blk_base + t0 + t1 + ((ex >> 32) & 0xFFFFull));
  slab(b0 + 3 * NL_SLAB, raw[3], ev[3], blk_base + t0 + t1 + t2 + ((ex >> 48) & 0xFFFFull));
}

// Builder-API records (Concurrent_Kmer_Partial_Counter::
// default_count_and_process, concurrent_kmer_counter.hpp:2018-2082): from
// the current position the scan finds the next skip symbol ('>'), skips to
// the delimiter ('\n') that ends the header, and takes read_len bytes from
// there (fewer at the end of the buffer) as the read; the next scan starts
// at the read's first byte.  The records are therefore the first skip symbol
// after every delimiter (and the first one of the buffer): with ev = the
// positions of all delimiter and skip bytes in order, event e is a record
// when it is a skip symbol and event e-1 is not.  A record whose header has
// no delimiter ends the scan; a record is counted when its read holds >=
// klen bytes.  The two symbols differ (checked on the host).
struct BldSyms {
  uint8_t delim, skip;
};

__device__ __forceinline__ bool bld_record(const uint8_t* fa, uint64_t n, const ull* ev, uint64_t m,
                                           uint64_t e, int klen, uint64_t read_len, BldSyms sy,
                                           uint64_t* rb, uint64_t* rl) {
  if (fa[ev[e]] != sy.skip || (e > 0 && fa[ev[e - 1]] == sy.skip)) return false;
  uint64_t e2 = e + 1;
  while (e2 < m && fa[ev[e2]] != sy.delim) e2++;
  if (e2 >= m) return false;
  *rb = ev[e2] + 1;
  *rl = read_len < n - *rb ? read_len : n - *rb;
  return *rl >= (uint64_t)klen;
}

__global__ void k_bld_flags(const uint8_t* fa, uint64_t n, const ull* ev, uint64_t m, int klen,
                            uint64_t read_len, BldSyms sy, uint32_t* flag) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  uint64_t rb, rl;
  flag[e] = bld_record(fa, n, ev, m, e, klen, read_len, sy, &rb, &rl) ? 1u : 0u;
}

__global__ void k_bld_write(const uint8_t* fa, uint64_t n, const ull* ev, uint64_t m, int klen,
                            uint64_t read_len, BldSyms sy, const ull* idx, ull* rstart,
                            uint32_t* rlen, uint32_t* kc) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  uint64_t rb, rl;
  if (!bld_record(fa, n, ev, m, e, klen, read_len, sy, &rb, &rl)) return;
  const ull r = idx[e];
  rstart[r] = rb;
  rlen[r] = (uint32_t)rl;
  kc[r] = (uint32_t)(rl - (uint64_t)klen + 1);
}

// Window records of the implicit builder counters
// (Concurrent_Kmer_Partial_Counter::concurrent_count_and_process_omp_implicit_v0
// / _v1, src/concurrent_kmer_counter.hpp:910-1022,1363-1578): the records of
// bld_record (same scan), every one counted, whatever its length.  flag[e] =
// event e starts a record; bad[0] = a read shorter than read_len (cut at the
// end of the buffer), bad[1] = a read shorter than klen.
__global__ void k_win_flags(const uint8_t* fa, uint64_t n, const ull* ev, uint64_t m, int klen,
                            uint64_t read_len, BldSyms sy, uint32_t* flag, uint32_t* bad) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  uint64_t rb, rl;
  const bool rec = bld_record(fa, n, ev, m, e, 0, read_len, sy, &rb, &rl);
  flag[e] = rec ? 1u : 0u;
  if (rec && rl != read_len) atomicOr(&bad[0], 1u);
  if (rec && rl < (uint64_t)klen) atomicOr(&bad[1], 1u);
}

// Writes the window records: start, length and windows (len - klen + 1, or 0).
__global__ void k_win_write(const uint8_t* fa, uint64_t n, const ull* ev, uint64_t m, int klen,
                            uint64_t read_len, BldSyms sy, const ull* idx, ull* rstart, uint32_t* rlen,
                            uint32_t* wc) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  uint64_t rb, rl;
  if (!bld_record(fa, n, ev, m, e, 0, read_len, sy, &rb, &rl)) return;
  const ull r = idx[e];
  rstart[r] = rb;
  rlen[r] = (uint32_t)rl;
  wc[r] = rl >= (uint64_t)klen ? (uint32_t)(rl - (uint64_t)klen + 1) : 0u;
}

// The byte values present in the reads: bit v of present[v / 32].
__global__ void k_win_alpha(const uint8_t* fa, const ull* rstart, const uint32_t* rlen, uint64_t R,
                            uint32_t* present) {
  __shared__ uint32_t p[8];
  if (threadIdx.x < 8) p[threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.y + threadIdx.y; r < R; r += (uint64_t)gridDim.x * blockDim.y) {
    const ull s = rstart[r];
    const uint32_t len = rlen[r];
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
      const uint32_t v = fa[s + i];
      atomicOr(&p[v >> 5], 1u << (v & 31));
    }
  }
  __syncthreads();
  if (threadIdx.x < 8 && threadIdx.y == 0 && p[threadIdx.x]) atomicOr(&present[threadIdx.x], p[threadIdx.x]);
}

// Window keys: the klen bytes of each window as codes of `bits` bits (code =
// rank of the byte value among the values present, so equal windows, and
// only those, get equal keys), first byte highest.  strncmp compares up to the
// first NUL: with NUL present (code 0) every code after a window's first NUL
// is 0.  One thread per read, windows in read order at woff[r].
template <class Key>
__global__ void k_win_keys(const uint8_t* fa, const ull* rstart, const uint32_t* rlen, const ull* woff,
                           uint64_t R, int klen, int bits, const uint8_t* lut, int has_nul, Key* S) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const ull s = rstart[r];
  const uint32_t len = rlen[r];
  if (len < (uint32_t)klen) return;
  const uint32_t nw = len - (uint32_t)klen + 1;
  const int kb = klen * bits;
  const Key mask = kb >= (int)(8 * sizeof(Key)) ? ~Key(0) : ((Key(1) << kb) - 1);
  Key* out = S + woff[r];
  if (!has_nul) {
    Key km = 0;
    for (uint32_t i = 0; i < len; i++) {
      km = ((km << bits) | (Key)lut[fa[s + i]]) & mask;
      if (i + 1 >= (uint32_t)klen) out[i + 1 - (uint32_t)klen] = km;
    }
    return;
  }
  for (uint32_t w = 0; w < nw; w++) {
    Key km = 0;
    bool nul = false;
    for (int j = 0; j < klen; j++) {
      const uint8_t v = fa[s + w + j];
      const uint32_t code = nul ? 0u : (uint32_t)lut[v];
      nul = nul || v == 0;
      km = (km << bits) | (Key)code;
    }
    out[w] = km;
  }
}

// Records (kmer_counting.cpp:2920-2945).  Candidate r uses header newline
// nl[j0+2r] and sequence newline nl[j0+2r+1] (or the end of the buffer).
__global__ void k_reads(uint64_t n, const ull* nl, uint64_t m, uint64_t p0,
                        int K, uint64_t ncand, ull* rstart, uint32_t* rlen, uint32_t* kc,
                        ull* first_unproc, ull* first_err, ull* too_long) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // j0 = first newline at or after p0 (one search per block)
  __shared__ uint64_t sj0;
  if (threadIdx.x == 0) {
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if ((nl[mid] & NL_POS_MASK) < p0) lo = mid + 1; else hi = mid;
    }
    sj0 = lo;
  }
  __syncthreads();
  if (r >= ncand) return;
  const uint64_t j0 = sj0;
  const uint64_t h = j0 + 2 * r;
  bool processed = false;
  if (p0 < n && h < m) {
    const uint64_t hdr_nl = nl[h] & NL_POS_MASK;
    processed = hdr_nl + 1 + (uint64_t)K <= n;
    if (processed) {
      const uint64_t s = hdr_nl + 1;
      const uint64_t e = (h + 1 < m) ? nl[h + 1] & NL_POS_MASK : n;
      const uint64_t len = e - s;
      rstart[r] = s;
      if (len > 0xFFFFFFF0ull) {
        atomicMin(too_long, (ull)r);
        rlen[r] = 0;
        kc[r] = 0;
      } else {
        rlen[r] = (uint32_t)len;
        kc[r] = len >= (uint64_t)K ? (uint32_t)(len - K + 1) : 0u;
      }
    }
  }
  if (!processed) {
    rstart[r] = 0;
    rlen[r] = 0;
    kc[r] = 0;
    atomicMin(first_unproc, (ull)r);
  }
  // assert(ptr[p] == '>') at the start of record r >= 1 when it is reached
  if (r >= 1 && p0 < n) {
    const uint64_t hp = j0 + 2 * r - 1;
    if (hp < m && (nl[hp] >> NL_FLAG_SHIFT) == 0) atomicMin(first_err, (ull)r);
  }
}

// ---------------------------------------------------------------------------
// Per-read rolling k-mer parse (one thread per read).

__device__ __forceinline__ uint64_t load_word8(const uint8_t* fa, uint64_t n, uint64_t a) {
  if (a + 8 <= n) return *reinterpret_cast<const uint64_t*>(fa + a);
  uint64_t w = 0;
  for (int i = 0; i < 8; i++)
    if (a + i < n) w |= (uint64_t)fa[a + i] << (8 * i);
  return w;
}

// Base encodings of the two counting front ends.
//   ENC_MAIN:    char_to_el, (c >> 1) & 7 (distribute_kmers.h:784-786): the
//                3-bit code ORs into the previous base's low bit, and the
//                word is masked to 2k bits (kmer_shift, :798-805).
//   ENC_BUILDER: symbol_to_kmer (concurrent_kmer_counter.hpp:2097-2110):
//                A C T G -> 0 1 2 3, every other byte -> 0; the word is not
//                masked and keeps the last 4*sizeof(kmer_t) bases.
constexpr int ENC_MAIN = 0;
constexpr int ENC_BUILDER = 1;

template <int ENC>
__device__ __forceinline__ uint32_t base_code(uint32_t c) {
  if (ENC == ENC_MAIN) return (c >> 1) & 7u;
  const bool acgt = c == 'A' || c == 'C' || c == 'G' || c == 'T';
  return acgt ? (c >> 1) & 3u : 0u;
}

// ---------------------------------------------------------------------------
// Partition (multisplit) into buckets by the top bits of the key.
//   level 1  k_l1_hist / k_l1_scatter: fused with the parse of the reads, up
//            to 10 bits; each key goes straight to its slot
//   level 2+ k_ms_hist / k_ms_scatter: from the previous level's array, at
//            most 8 bits so that every tile holds tens of keys per digit:
//            each tile is sorted by digit in LDS and written as runs

constexpr int L1_NT = 512;  // level-1 histogram blocks
constexpr int MS_NT = 512;
constexpr int MS_P_NT = 512;  // threads of a persistent multisplit block (one tile at a time)
constexpr int L1_MAX_BITS = PK_TUNING_L1_BITS_MAX;  // level-1 digit width (direct scatter)
constexpr int L1_DEFAULT_BITS = 8;  // wider digits keep more partial lines open
constexpr int MS_MAX_BITS = 8;   // further levels (tile split into runs)
constexpr int MS_TIMED_LEVELS = 4;  // further levels with their own timing events
static_assert(sizeof(pk_timing::ms_level_form) / sizeof(int32_t) == MS_TIMED_LEVELS, "per-level timing slots");
constexpr uint64_t L2_FOLD_MAX = 8192;  // (bucket, digit) counters of a folded level-2 histogram

template <class Key>
struct TileCfg {
  static constexpr int Q = sizeof(Key) == 8 ? 16 : 8;  // keys per thread per tile
};

// LDS of a tile split by digit: a histogram and its exclusive scan, a
// counting-sort scatter of the keys into T1, then coalesced stores: sorted
// slot p, of digit d (recomputed from the key), goes to out[base[d] -
// toff[d] + p] (delta[d] = base[d] - toff[d]; base[d] stays in the register
// of thread d): about 68 KB for 8192 one-word or 4096 two-word keys.
template <class Key, int NT>
struct TileLds {
  Key T1[NT * TileCfg<Key>::Q];
  uint32_t h[256];
  uint32_t tcur[256];
  long long delta[256];
};

// Level-1 producer.  A block owns rpb consecutive reads and walks them in
// chunks: a chunk is the longest run of <= NT consecutive reads (NT threads
// per block) whose bytes fit in the LDS staging buffer; thread t then walks
// read t of the chunk from LDS with the rolling update (a read too long for
// the buffer is walked from global memory, alone in its chunk).  Records are in file
// order, so a chunk's bytes are one contiguous range.  The buffer holds the
// base code of every byte, not the byte: bits 0-1 of the code in plane A
// (32 bytes per word) and, on the main path, bit 2 in plane B (32 bytes per
// word), 3/8 of the raw size.  Its size (sb bytes, in dynamic LDS) is chosen
// per input so that NT typical records fit.
constexpr int L1_SB_MIN = 8192, L1_SB_MAX = 98304;  // staged bytes per chunk

// Level-1 scatter blocks: 1024 threads on the main path (two blocks per CU,
// <= 60 KB of LDS and <= 64 VGPRs each; fewer blocks per CU keep fewer
// partly written lines open), 512 for the rank-share and builder walks,
// which need more registers.  With one-word keys a scatter block takes the
// reads of two histogram blocks (k = 32 bacterial 1.27 -> 1.22 ms; the
// histogram itself is faster at 512 threads over half as many reads); with
// two-word keys one (k = 64 bacterial 1.50 -> 1.34 ms against 512 threads).
template <class Key, int ENC, bool OWN>
struct L1Scatter {
  static constexpr bool main = ENC == ENC_MAIN && !OWN;
  static constexpr int NT = main ? 1024 : 512;
  static constexpr int SB_MAX = main ? 131072 : L1_SB_MAX;
  // histogram blocks per scatter block (their output runs per digit are
  // adjacent in the digit-major scan)
  static constexpr int GF = main && sizeof(Key) == 8 ? 2 : 1;
};

// Level-2 histogram folded into the level-1 histogram (large inputs with
// three or more partition levels): the level-2 "segments" are runs of the
// level-1 output, segment (a, g) = the keys of level-1 bucket a that the
// histogram blocks [g * spb, (g + 1) * spb) produced (spb a multiple of the
// histogram blocks per scatter block, so a segment is whole scatter-block
// runs, contiguous in the digit-major layout); the level-1 histogram kernel
// counts them per second digit s into H2[(a, s) * nsb + g], so no pass over
// the written keys is needed to count them.
struct L2Fold {
  int d2;         // second-digit bits (0: no fold)
  int spb;        // histogram blocks per segment
  uint64_t nsb;   // segments per level-1 bucket
  uint32_t* H2;   // (2^(d1+d2)) x nsb counts, zeroed before the launch
};

struct ChunkLds {
  ull base;  // first staged byte (32-byte aligned)
  int m;     // reads in the chunk
  int lds;   // bytes are staged
};

struct Stage {
  ull* A;       // code bits 0-1 of byte 32 w + i at bits 2i, 2i+1 of A[w]
  uint32_t* B;  // code bit 2 of byte 32 w + i at bit i of B[w] (main path)
  int sb;       // capacity in bytes (a multiple of 32)
};

// Bytes of dynamic LDS of a level-1 kernel: the two planes, then nd digit
// entries of dig_bytes each.
__host__ __device__ constexpr uint32_t l1_lds_bytes(int sb, int nd, int dig_bytes) {
  return (uint32_t)(sb / 32) * 12u + (uint32_t)nd * (uint32_t)dig_bytes;
}

// Sets up the chunk starting at read c0 (< r1) and stages its codes.
template <int ENC, int NT>
__device__ __forceinline__ void stage_chunk(ChunkLds& C, Stage sg, const uint8_t* fa, uint64_t n,
                                            const ull* rstart, const uint32_t* rlen, uint64_t c0,
                                            uint64_t r1) {
  const uint64_t r = c0 + threadIdx.x;
  const ull first = rstart[c0] & ~31ull;
  int fits = 0;
  if (r < r1 && (int)threadIdx.x < NT) fits = rstart[r] + rlen[r] - first <= (ull)sg.sb;
  // records ascend in the file: the reads that fit are a prefix
  const int m = (int)block_sum<NT>((ull)fits);
  if (threadIdx.x == 0) {
    C.base = first;
    C.m = m > 0 ? m : 1;
    C.lds = m > 0;
  }
  __syncthreads();
  if (C.lds) {
    const uint64_t last = c0 + C.m - 1;
    const ull end = rstart[last] + rlen[last];
    const ull nw = (end - first + 31) / 32;
    for (ull i = threadIdx.x; i < nw; i += NT) {
      ull a = 0;
      uint32_t b = 0;
#pragma unroll
      for (int h = 0; h < 4; h++) {
        const ull w = load_word8(fa, n, first + 32 * i + 8 * h);
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const uint32_t el = base_code<ENC>((uint32_t)(w >> (8 * k)) & 0xFFu);
          a |= (ull)(el & 3u) << (16 * h + 2 * k);
          b |= (el >> 2) << (8 * h + k);
        }
      }
      sg.A[i] = a;
      if (ENC == ENC_MAIN) sg.B[i] = b;
    }
  }
  __syncthreads();
}

// Walks up to Q more k-mers of the thread's read (state s, i, km): emits
// through emit(q, key); returns how many.
// Rank-share routing of the reference's Sliding_window (kmer_counting.cpp:
// 2950-3050): a k-mer belongs to rank hash31(m) % P (retrieve_proc_id,
// distribute_kmers.cpp:95-105; hash31, distribute_kmers.h:856-873) of its
// leftmost minimum-frequency L-mer m (L = 8, frequencies of the whole input).
// The minimum is tracked as the window slides: a new L-mer replaces it only
// when strictly rarer, and when it leaves the window every L-mer of the
// current k-mer is re-read from the packed k-mer (recalculate_min_lmer,
// kmer_counting.cpp:355-388).  The rolling L-mer uses the 3-bit code like
// the k-mer; the re-read extracts 2-bit fields of the k-mer.
constexpr int LMER_L = 8;
constexpr uint32_t LMER_BINS = 1u << (2 * LMER_L);

struct OwnCfg {
  const ull* freq;  // LMER_BINS L-mer frequencies (min-L-mer owner)
  int P, rank;
  int by_key;       // owner of the whole k-mer instead (the -p counters)
};

struct OwnState {
  uint32_t lm, min_lmer;
  ull min_f;
  int min_pos;
};

// retrieve_proc_id of a k-mer word (src/distribute_kmers.cpp:95-120): hash31
// (includes/distribute_kmers.h:856-873: signed 64-bit a * x + b, arithmetic
// shift) when kmer_t is 64 bits, uhash31 (:836-854: the same in unsigned
// 128-bit arithmetic) when it is 128 bits; then (int) % P.  Both wrap as the
// reference's builds do.
__device__ __forceinline__ int key_owner(uint64_t x, int P) {
  const long long v = (long long)(68111ull * x + 105929ull);
  const long long h = (long long)((ull)(v >> 31) + (ull)v) & 2147483647LL;
  return (int)h % P;
}
__device__ __forceinline__ int key_owner(u128 x, int P) {
  u128 r = (u128)68111 * x + (u128)105929;
  r = ((r >> 31) + r) & (u128)2147483647;
  return (int)(long long)r % P;
}

__device__ __forceinline__ int lmer_owner(uint32_t lmer, int P) {
  const long long v = 68111LL * (long long)lmer + 105929LL;  // hasha * x + hashb
  const long long h = ((v >> 31) + v) & 2147483647LL;
  return (int)h % P;
}

// One byte of a read (index idx, 3-bit code el; km already updated).
// Returns whether the k-mer ending here, if any, is owned.
template <class Key, bool OWN>
__device__ __forceinline__ bool own_step(OwnState& o, const OwnCfg& cf, Key km, uint32_t el, uint32_t idx,
                                         int K) {
  if (!OWN) return true;
  // the -p counters own a k-mer by a hash of the whole word
  // (kmer_counting.cpp:1949-1950)
  if (cf.by_key) return idx >= (uint32_t)(K - 1) && key_owner(km, cf.P) == cf.rank;
  o.lm = ((o.lm << 2) | el) & (LMER_BINS - 1);
  if (idx < (uint32_t)(K - 1)) {
    if (idx >= (uint32_t)(LMER_L - 1)) {
      const ull f = cf.freq[o.lm];
      if (idx == (uint32_t)(LMER_L - 1)) {
        o.min_lmer = o.lm;
        o.min_f = f;
      } else if (f < o.min_f) {
        o.min_lmer = o.lm;
        o.min_f = f;
        o.min_pos = (int)idx - (LMER_L - 1);
      }
    }
    return false;
  }
  const ull f = cf.freq[o.lm];
  if (o.min_pos < 0) {
    for (int kk = 0; kk <= K - LMER_L; kk++) {
      const uint32_t lo = (uint32_t)(km >> (2 * (K - kk - LMER_L))) & (LMER_BINS - 1);
      const ull g = cf.freq[lo];
      if (kk == 0 || g < o.min_f) {
        o.min_lmer = lo;
        o.min_f = g;
        o.min_pos = kk;
      }
    }
  }
  if (f < o.min_f) {
    o.min_lmer = o.lm;
    o.min_f = f;
    o.min_pos = K - LMER_L;
  }
  o.min_pos--;
  return lmer_owner(o.min_lmer, cf.P) == cf.rank;
}

// km rolls unmasked (bits above the key shift out of the word, and the
// key's own bits are the same either way): the mask is applied to the keys
// handed out, once per k-mer instead of once per base.
template <class Key, int ENC, bool OWN, class Emit>
__device__ __forceinline__ int walk_read(const ChunkLds& C, Stage sg, const uint8_t* fa, ull s, uint32_t len,
                                         uint32_t& i,
                                         Key& km, OwnState& os, const OwnCfg& cf, int K, Key mask, int Q,
                                         Emit&& emit) {
  int q = 0;
  if (C.lds) {
    // 32 staged codes at a time from the read's own position (two words
    // shifted together), so every lane runs the same count but on its last
    // piece (pieces ending at word boundaries differ per read and diverge)
    while (q < Q && i < len) {
      const ull off = s + i - C.base;
      const ull w = off >> 5;
      const uint32_t sh = (uint32_t)(off & 31);
      const uint32_t cnt = len - i < 32u ? len - i : 32u;
      ull a = sg.A[w] >> (2 * sh);
      uint32_t b = ENC == ENC_MAIN ? sg.B[w] >> sh : 0u;
      if (sh + cnt > 32u) {
        // the piece runs into the next word, which holds bases of this read
        a |= sg.A[w + 1] << (64 - 2 * sh);
        if (ENC == ENC_MAIN) b |= sg.B[w + 1] << (32 - sh);
      }
      for (uint32_t t = 0; t < cnt && q < Q; t++) {
        const uint32_t el = ((uint32_t)a & 3u) | ((b & 1u) << 2);
        km = (km << 2) | (Key)el;
        const bool own = own_step<Key, OWN>(os, cf, OWN ? km & mask : km, el, i, K);
        a >>= 2;
        b >>= 1;
        i++;
        if (i >= (uint32_t)K && own) emit(q++, km & mask);
      }
    }
    return q;
  }
  while (q < Q && i < len) {
    const uint32_t el = base_code<ENC>(fa[s + i]);
    km = (km << 2) | (Key)el;
    const bool own = own_step<Key, OWN>(os, cf, OWN ? km & mask : km, el, i, K);
    i++;
    if (i >= (uint32_t)K && own) emit(q++, km & mask);
  }
  return q;
}

// The histogram's digit of a main-path k-mer is its top d <= 13 key bits:
// bases 0..6 of the window (key slot j holds base j's 2-bit code at bits
// 2(K-1-j) and 2(K-1-j)+1, and the third code bit of base j+1 ORs into the
// low bit of slot j).  A 15-bit register rolled like the key, ((v << 2) |
// code) & 0x7FFF, holds slots 0..6 of the window that starts 6 bases back
// (bit 14, the spill of the bases before, is never read; bit 0 still lacks
// the next base's third bit, so d stays <= 13), and digit = (v >> (14 - d))
// & (2^d - 1).  The walk stops at the last window's
// seventh base: len - K + 7 bases per read instead of len, no K-bit key.
// For K >= 7 only.
template <class Emit>
__device__ __forceinline__ void walk_read_top(const ChunkLds& C, Stage sg, const uint8_t* fa, ull s, uint32_t len,
                                              int K, Emit&& emit) {
  if (len < (uint32_t)K) return;
  const uint32_t end = len - (uint32_t)K + 7;  // bases to read
  uint32_t v = 0, i = 0;
  if (C.lds) {
    // 32 bases at a time from the read's own position (two words shifted
    // together), so every lane runs the same count but on its last piece
    // (pieces ending at word boundaries differ per read and diverge)
    while (i < end) {
      const ull off = s + i - C.base;
      const ull w = off >> 5;
      const uint32_t sh = (uint32_t)(off & 31);
      const uint32_t cnt = end - i < 32u ? end - i : 32u;
      ull a = sg.A[w] >> (2 * sh);
      uint32_t b = sg.B[w] >> sh;
      if (sh + cnt > 32u) {
        // the piece runs into the next word, which holds bases of this read
        a |= sg.A[w + 1] << (64 - 2 * sh);
        b |= sg.B[w + 1] << (32 - sh);
      }
      for (uint32_t q = 0; q < cnt; q++, i++) {
        v = ((v << 2) | ((uint32_t)a & 3u) | ((b & 1u) << 2)) & 0x7FFFu;
        a >>= 2;
        b >>= 1;
        if (i >= 6) emit(v);
      }
    }
    return;
  }
  for (; i < end; i++) {
    v = ((v << 2) | base_code<ENC_MAIN>(fa[s + i])) & 0x7FFFu;
    if (i >= 6) emit(v);
  }
}

// L-mer frequencies (Sliding_window_l, kmer_counting.cpp:391-438): every
// L-mer of every record's read line, rolling 3-bit update masked to 2L bits.
__global__ void k_lmer_freq(const uint8_t* fa, const ull* rstart, const uint32_t* rlen, uint64_t R,
                            ull* freq) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const ull s = rstart[r];
    const uint32_t len = rlen[r];
    uint32_t lm = 0;
    for (uint32_t i = 0; i < len; i++) {
      lm = ((lm << 2) | base_code<ENC_MAIN>(fa[s + i])) & (LMER_BINS - 1);
      if (i >= (uint32_t)(LMER_L - 1)) atomicAdd(&freq[lm], 1ull);
    }
  }
}

// Level-1 histogram: digit = bits [KB - d1, KB) of each key (KB = key bits:
// 2K on the main path, the whole word for the builder encoding).  Block b's
// count of digit d goes to H[d * nblk + b] (digit-major), so an exclusive
// scan of H gives every block the output base of each of its digit runs.
template <class Key, int ENC, bool OWN>
__global__ __launch_bounds__(L1_NT) void k_l1_hist(const uint8_t* fa, uint64_t n, const ull* rstart,
                                                   const uint32_t* rlen, uint64_t R, int K, int KB,
                                                   Key mask, int d1, int rpb, int sb, OwnCfg cf,
                                                   uint32_t* H, L2Fold f2) {
  __shared__ ChunkLds C;
  // dynamic LDS: plane A, plane B, 2^d1 digit counts (then, with f2.d2 > 0,
  // 2^(d1 + d2) counts of the first two digits)
  extern __shared__ ull l1_dyn[];
  const Stage sg{l1_dyn, reinterpret_cast<uint32_t*>(l1_dyn + sb / 32), sb};
  uint32_t* h = sg.B + sb / 32;
  const int nb = 1 << d1;
  uint32_t* h2 = h + nb;
  const int nb2 = f2.d2 ? nb << f2.d2 : 0;
  for (int i = threadIdx.x; i < nb; i += L1_NT) h[i] = 0;
  for (int i = threadIdx.x; i < nb2; i += L1_NT) h2[i] = 0;
  const uint64_t r0 = (uint64_t)blockIdx.x * rpb;
  const uint64_t r1 = r0 + rpb < R ? r0 + rpb : R;
  const int bpos = KB - d1;
  // the main path's digits (<= 13 bits: the register's bit 0, the low bit of
  // slot 6, still lacks the next base's third code bit) come from a 15-bit
  // window of bases
  const bool top = ENC == ENC_MAIN && !OWN && K >= 7 && KB == 2 * K && d1 + f2.d2 <= 13;
  const int tsh = 14 - d1, tsh2 = 14 - d1 - f2.d2;
  const uint32_t tm = (1u << d1) - 1u, tm2 = (1u << (d1 + f2.d2)) - 1u;
  for (uint64_t c0 = r0; c0 < r1;) {
    stage_chunk<ENC, L1_NT>(C, sg, fa, n, rstart, rlen, c0, r1);
    const int m = C.m;
    if ((int)threadIdx.x < m) {
      const uint64_t r = c0 + threadIdx.x;
      if (top) {
        // main path: the digit from the window's first seven bases
        if (nb2) walk_read_top(C, sg, fa, rstart[r], rlen[r], K, [&](uint32_t v) { atomicAdd(&h2[(v >> tsh2) & tm2], 1u); });
        else walk_read_top(C, sg, fa, rstart[r], rlen[r], K, [&](uint32_t v) { atomicAdd(&h[(v >> tsh) & tm], 1u); });
      } else {
        uint32_t i = 0;
        Key km = 0;
        OwnState os{0, 0, 0, 0};
        walk_read<Key, ENC, OWN>(C, sg, fa, rstart[r], rlen[r], i, km, os, cf, K, mask, 0x7fffffff,
                                 [&](int, Key x) {
                                   // one LDS atomic per key: with the fold the
                                   // level-1 counts are sums of h2's rows
                                   if (nb2) atomicAdd(&h2[key_digit<Key>(x, bpos - f2.d2, d1 + f2.d2)], 1u);
                                   else atomicAdd(&h[key_digit<Key>(x, bpos, d1)], 1u);
                                 });
      }
    }
    c0 += m;
    __syncthreads();
  }
  if (nb2) {
    // h[a] = sum of h2[a, 0 .. 2^d2): one wave per bucket, lanes over s
    const int lane = (int)threadIdx.x & 63, per = 1 << f2.d2;
    for (int a = (int)threadIdx.x >> 6; a < nb; a += L1_NT / 64) {
      uint32_t v = 0;
      for (int j = lane; j < per; j += 64) v += h2[(a << f2.d2) + j];
      for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) h[a] = v;
    }
    __syncthreads();
    // (bucket, sub-digit, segment) counts of the level-2 segments
    for (int i = threadIdx.x; i < nb2; i += L1_NT)
      if (h2[i]) atomicAdd(&f2.H2[(ull)i * f2.nsb + blockIdx.x / f2.spb], h2[i]);
  }
  for (int i = threadIdx.x; i < nb; i += L1_NT) H[(ull)i * gridDim.x + blockIdx.x] = h[i];
}

// Level-1 scatter fused with the parse: the same chunked walk as k_l1_hist
// over the same reads per block, in blocks of L1Scatter::NT threads; every key
// takes the next slot of its digit from the block's LDS counter (atomicAdd's
// return value) and is stored at once at the block's base for that digit
// (scanned histogram O, same layout as H).  No tile buffer and no barrier
// inside the walk: the order of keys inside a bucket is arbitrary, which the
// bucket count does not depend on.  Each block keeps a partly written line
// open per digit, so fewer, larger blocks per CU write fewer partial lines.
template <class Key, int ENC, bool OWN>
__global__ __launch_bounds__((L1Scatter<Key, ENC, OWN>::NT)) void k_l1_scatter(const uint8_t* fa, uint64_t n,
                                                      const ull* rstart, const uint32_t* rlen,
                                                      uint64_t R, int K, int KB, Key mask, int d1,
                                                      int rpb, int sb, OwnCfg cf, const ull* O,
                                                      uint64_t nblk_h, Key* S1) {
  constexpr int NT = L1Scatter<Key, ENC, OWN>::NT;
  constexpr int GF = L1Scatter<Key, ENC, OWN>::GF;
  __shared__ ChunkLds C;
  // dynamic LDS: plane A, 2^d1 output cursors, plane B (then 2^d1 words
  // the cursors no longer need: each k-mer takes its slot from one 64-bit
  // atomic on its digit's cursor)
  extern __shared__ ull l1_dyn[];
  const int nb = 1 << d1;
  ull* cur = l1_dyn + sb / 32;
  const Stage sg{l1_dyn, reinterpret_cast<uint32_t*>(cur + nb), sb};
  for (int i = threadIdx.x; i < nb; i += NT) cur[i] = O[(ull)i * nblk_h + (ull)blockIdx.x * GF];
  const uint64_t r0 = (uint64_t)blockIdx.x * GF * rpb;
  const uint64_t r1 = r0 + (uint64_t)GF * rpb < R ? r0 + (uint64_t)GF * rpb : R;
  const int bpos = KB - d1;
  // stage_chunk's barriers order the set-up above before the first walk
  for (uint64_t c0 = r0; c0 < r1;) {
    stage_chunk<ENC, NT>(C, sg, fa, n, rstart, rlen, c0, r1);
    const int m = C.m;
    if ((int)threadIdx.x < m) {
      const uint64_t r = c0 + threadIdx.x;
      uint32_t i = 0;
      Key km = 0;
      OwnState os{0, 0, 0, 0};
      walk_read<Key, ENC, OWN>(C, sg, fa, rstart[r], rlen[r], i, km, os, cf, K, mask, 0x7fffffff,
                               [&](int, Key x) {
                                 const uint32_t d = key_digit<Key>(x, bpos, d1);
                                 S1[atomicAdd(&cur[d], 1ull)] = x;
                               });
    }
    c0 += m;
    __syncthreads();
  }
}

// bucket offsets of a level from its scanned digit-major histogram:
// boff[d] = O[d * stride_of(d)] ... for level 1 stride = nblk.
__global__ void k_l1_boff(const ull* O, int nb, uint64_t nblk, ull total, ull* boff) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < nb) boff[d] = O[(ull)d * nblk];
  if (d == nb) boff[nb] = total;
}

// tiles of a level: tp[a+1] = tp[a] + ceil(n_a / tile) for a < nb (one block)
__global__ __launch_bounds__(1024) void k_tile_prefix(const ull* boff, int nb, int tile, ull* tp) {
  __shared__ ull tmp[1025];
  ull acc = 0;
  if (threadIdx.x == 0) tp[0] = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int a = base + threadIdx.x;
    const ull v = a < nb ? (boff[a + 1] - boff[a] + tile - 1) / tile : 0ull;
    const ull ex = block_excl_scan<1024>(v);
    if (a < nb) tp[a + 1] = acc + ex + v;
    tmp[threadIdx.x] = ex + v;
    __syncthreads();
    acc += tmp[1023];
    __syncthreads();
  }
}

__device__ __forceinline__ int tile_owner(const ull* tp, int nb, ull blk) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tp[mid] <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Level >= 2: bucket a of the previous level (nb buckets, offsets boff,
// tiles tp) is split by dbits more bits at dpos.  Tile t of bucket a writes
// its count of sub-digit s to H[tp[a] << dbits + s * ntiles(a) + t]: in
// (a, s, t) order, so the exclusive scan of H holds every tile's output base
// per sub-digit and H's scan at t = 0 is the start of sub-bucket (a, s).
__device__ __forceinline__ ull ms_index(const ull* tp, int a, int s, ull t, int dbits) {
  return (tp[a] << dbits) + (ull)s * (tp[a + 1] - tp[a]) + t;
}

template <class Key>
__global__ __launch_bounds__(MS_NT) void k_ms_hist(const Key* Sin, const ull* boff, const ull* tp, int nb,
                                                   int dpos, int dbits, uint32_t* H) {
  constexpr int TQ = MS_NT * TileCfg<Key>::Q;
  __shared__ uint32_t h[256];
  const ull blk = blockIdx.x;
  if (blk >= tp[nb]) return;
  const int a = tile_owner(tp, nb, blk);
  const ull t = blk - tp[a];
  const ull b0 = boff[a] + t * TQ;
  const ull b1 = b0 + TQ < boff[a + 1] ? b0 + TQ : boff[a + 1];
  const int ns = 1 << dbits;
  for (int i = threadIdx.x; i < ns; i += MS_NT) h[i] = 0;
  __syncthreads();
  for (ull i = b0 + threadIdx.x; i < b1; i += MS_NT) atomicAdd(&h[key_digit<Key>(Sin[i], dpos, dbits)], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < ns; i += MS_NT) H[ms_index(tp, a, i, t, dbits)] = h[i];
}

// boff2[a << dbits | s] = O[ms_index(a, s, 0)] for every (a, s); boff2[end] = N
__global__ void k_ms_boff(const ull* O, const ull* tp, int nb, int dbits, ull total, ull* boff2) {
  const ull g = (ull)blockIdx.x * blockDim.x + threadIdx.x;
  const ull n2 = (ull)nb << dbits;
  if (g < n2) {
    const int a = (int)(g >> dbits);
    const int s = (int)(g & ((1u << dbits) - 1));
    boff2[g] = (tp[a + 1] > tp[a]) ? O[ms_index(tp, a, s, 0, dbits)] : O[tp[a] << dbits];
  }
  if (g == n2) boff2[n2] = total;
}

// Tile geometry: bucket a (advanced from the previous tile's, tiles being
// visited in order), first key b0 and key count m of tile g.
__device__ __forceinline__ void ms_tile(const ull* boff, const ull* tp, ull g, int TQ, int& a, ull& t,
                                        ull& b0, int& m) {
  while (tp[a + 1] <= g) a++;
  t = g - tp[a];
  b0 = boff[a] + t * TQ;
  const ull e = boff[a + 1];
  m = (int)((b0 + TQ < e ? b0 + TQ : e) - b0);
}

template <class Key>
__global__ __launch_bounds__(MS_NT) void k_ms_scatter(const Key* Sin, const ull* boff, const ull* tp,
                                                      int nb, int dpos, int dbits, const ull* O,
                                                      Key* Sout) {
  constexpr int Q = TileCfg<Key>::Q;
  constexpr int TQ = MS_NT * TileCfg<Key>::Q;
  __shared__ TileLds<Key, MS_NT> L;
  const ull blk = blockIdx.x;
  if (blk >= tp[nb]) return;
  const int a = tile_owner(tp, nb, blk);
  const ull t = blk - tp[a];
  const ull b0 = boff[a] + t * TQ;
  const ull b1 = b0 + TQ < boff[a + 1] ? b0 + TQ : boff[a + 1];
  const int m = (int)(b1 - b0);
  const int ns = 1 << dbits;
  // sub-digit tid's output base (ns <= 256 <= MS_NT)
  const ull bse = (int)threadIdx.x < ns ? O[ms_index(tp, a, threadIdx.x, t, dbits)] : 0ull;
  if ((int)threadIdx.x < ns) L.h[threadIdx.x] = 0;
  // keys tid, tid + MS_NT, ... of the tile in registers (coalesced loads)
  Key kr[Q];
#pragma unroll
  for (int j = 0; j < Q; j++) {
    const int idx = (int)threadIdx.x + j * MS_NT;
    kr[j] = idx < m ? Sin[b0 + idx] : Key(0);
  }
  const int np = m > (int)threadIdx.x ? (m - (int)threadIdx.x - 1) / MS_NT + 1 : 0;
  __syncthreads();
  // the histogram's atomic returns each key's rank inside its digit (two
  // 16-bit ranks per register), so the sort needs no second atomic pass
  uint32_t rk[(Q + 1) / 2] = {};
#pragma unroll
  for (int j = 0; j < Q; j++)
    if (j < np) rk[j / 2] |= atomicAdd(&L.h[key_digit<Key>(kr[j], dpos, dbits)], 1u) << (16 * (j & 1));
  __syncthreads();
  const uint32_t v = (int)threadIdx.x < ns ? L.h[threadIdx.x] : 0u;
  const ull ex = block_excl_scan<MS_NT>(v);
  if ((int)threadIdx.x < ns) {
    L.tcur[threadIdx.x] = (uint32_t)ex;
    L.delta[threadIdx.x] = (long long)bse - (long long)ex;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < Q; j++) {
    if (j < np) {
      const uint32_t d = key_digit<Key>(kr[j], dpos, dbits);
      L.T1[L.tcur[d] + ((rk[j / 2] >> (16 * (j & 1))) & 0xFFFFu)] = kr[j];
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < m; p += MS_NT) {
    const Key v = L.T1[p];
    Sout[L.delta[key_digit<Key>(v, dpos, dbits)] + p] = v;
  }
}

// Persistent form for long levels: block i takes the tiles
// [i * per, (i + 1) * per) in order and loads the next tile's keys while it
// sorts the current one (at most 128 VGPRs: two blocks of 8 waves per CU).
// Consecutive tiles of a bucket write adjacent runs of each sub-digit, so one
// CU completes the lines at their joins.  The output base of sub-digit s for
// tile t + 1 of a bucket is tile t's base plus tile t's count of s (the scan
// of H in (bucket, digit, tile) order), so bases are read from O only for a
// block's first tile in each bucket: one line per digit per bucket instead of
// per tile.
template <class Key, int NT>
__global__ __launch_bounds__(NT, NT >= 1024 ? 8 : 4) void k_ms_scatter_p(const Key* Sin, const ull* boff, const ull* tp,
                                                           int nb, int dpos, int dbits, const ull* O,
                                                           Key* Sout) {
  // NT threads share one tile of TQ keys (the level's tile), Q per thread
  constexpr int TQ = MS_NT * TileCfg<Key>::Q;
  constexpr int Q = TQ / NT;
  __shared__ TileLds<Key, MS_NT> L;
  const ull ntl = tp[nb];
  const ull per = (ntl + gridDim.x - 1) / gridDim.x;
  const ull g0 = (ull)blockIdx.x * per;
  const ull g1 = g0 + per < ntl ? g0 + per : ntl;
  if (g0 >= g1) return;
  const int ns = 1 << dbits;
  const int tid = (int)threadIdx.x;
  // the current tile's keys (tid, tid + NT, ...) and this thread's
  // sub-digit base (tid < ns)
  int a = tile_owner(tp, nb, g0);
  ull t, b0;
  int m;
  ms_tile(boff, tp, g0, TQ, a, t, b0, m);
  Key kr[Q];
#pragma unroll
  for (int j = 0; j < Q; j++) {
    const int idx = tid + j * NT;
    kr[j] = idx < m ? Sin[b0 + idx] : Key(0);
  }
  ull bs = tid < ns ? O[ms_index(tp, a, tid, t, dbits)] : 0ull;
  for (ull g = g0; g < g1; g++) {
    Key kn[Q] = {};
    ull bn = 0, tn = 0, b0n = 0;
    int an = a, mn = 0;
    if (g + 1 < g1) {
      ms_tile(boff, tp, g + 1, TQ, an, tn, b0n, mn);
#pragma unroll
      for (int j = 0; j < Q; j++) {
        const int idx = tid + j * NT;
        kn[j] = idx < mn ? Sin[b0n + idx] : Key(0);
      }
      // a new bucket: its first base from the scan
      if (tid < ns && an != a) bn = O[ms_index(tp, an, tid, tn, dbits)];
    }
    if (tid < ns) L.h[tid] = 0;
    const int np = m > tid ? (m - tid - 1) / NT + 1 : 0;
    __syncthreads();
    // ranks inside each digit from the histogram's atomics (see k_ms_scatter)
    uint32_t rk[(Q + 1) / 2] = {};
#pragma unroll
    for (int j = 0; j < Q; j++)
      if (j < np) rk[j / 2] |= atomicAdd(&L.h[key_digit<Key>(kr[j], dpos, dbits)], 1u) << (16 * (j & 1));
    __syncthreads();
    const uint32_t v = tid < ns ? L.h[tid] : 0u;
    // the next tile of the same bucket starts its runs where this one's end
    if (an == a) bn = bs + v;
    const ull ex = block_excl_scan<NT>(v);
    if (tid < ns) {
      L.tcur[tid] = (uint32_t)ex;
      L.delta[tid] = (long long)bs - (long long)ex;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < Q; j++) {
      if (j < np) {
        const uint32_t d = key_digit<Key>(kr[j], dpos, dbits);
        L.T1[L.tcur[d] + ((rk[j / 2] >> (16 * (j & 1))) & 0xFFFFu)] = kr[j];
      }
    }
    __syncthreads();
    for (int p = tid; p < m; p += NT) {
      const Key v = L.T1[p];
      Sout[L.delta[key_digit<Key>(v, dpos, dbits)] + p] = v;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < Q; j++) kr[j] = kn[j];
    bs = bn;
    a = an;
    m = mn;
  }
}

// Level 2 over the segments of the folded histogram (L2Fold): segment j =
// (a, g), a = j / nsb, holds [O1[a * nblk + g * spb], O1[a * nblk +
// min((g + 1) * spb, nblk)]) of the level-1 array; its run of sub-digit s
// starts at O2T[(j << dbits) + s] (the scan of H2, transposed so that a
// segment's bases are one contiguous run).  Block i takes
// the segments that start in its 1/G share of the keys and sorts them tile by
// tile (TQ keys) as k_ms_scatter_p does, the next tile's keys loaded ahead;
// within a segment each tile's runs start where the previous tile's ended.
__device__ __forceinline__ void seg_bounds(const ull* O1, uint64_t nblk, int spb, uint64_t nsb, ull j, ull& st,
                                           ull& en) {
  const ull a = j / nsb, sb = j - a * nsb;
  const ull lo = sb * (ull)spb;
  const ull hi = lo + (ull)spb < nblk ? lo + (ull)spb : nblk;
  st = O1[a * nblk + lo];
  en = O1[a * nblk + hi];
}

__device__ __forceinline__ ull seg_first_at(const ull* O1, uint64_t nblk, int spb, uint64_t nsb, ull nseg, ull k) {
  ull lo = 0, hi = nseg;
  while (lo < hi) {
    const ull mid = (lo + hi) >> 1;
    ull st, en;
    seg_bounds(O1, nblk, spb, nsb, mid, st, en);
    if (st < k) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <class Key, int NT>
__global__ __launch_bounds__(NT, 4) void k_ms_scatter_seg(const Key* Sin, const ull* O1, uint64_t nblk, int spb,
                                                          uint64_t nsb, int nb1, ull total, int dpos, int dbits,
                                                          const ull* O2T, Key* Sout) {
  constexpr int TQ = MS_NT * TileCfg<Key>::Q;
  constexpr int Q = TQ / NT;
  __shared__ TileLds<Key, MS_NT> L;
  const ull nseg = (ull)nb1 * nsb;
  const ull k0 = total * (ull)blockIdx.x / gridDim.x;
  const ull k1 = total * ((ull)blockIdx.x + 1) / gridDim.x;
  ull j = seg_first_at(O1, nblk, spb, nsb, nseg, k0);
  const ull j1 = blockIdx.x + 1 == gridDim.x ? nseg : seg_first_at(O1, nblk, spb, nsb, nseg, k1);
  const int ns = 1 << dbits;
  const int tid = (int)threadIdx.x;
  // first non-empty segment
  ull st = 0, en = 0;
  while (j < j1) {
    seg_bounds(O1, nblk, spb, nsb, j, st, en);
    if (en > st) break;
    j++;
  }
  if (j >= j1) return;
  auto base_of = [&](ull jj) -> ull { return O2T[(jj << dbits) + (ull)tid]; };
  ull pos = st;
  int m = (int)(en - pos < (ull)TQ ? en - pos : (ull)TQ);
  Key kr[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int idx = tid + q * NT;
    kr[q] = idx < m ? Sin[pos + idx] : Key(0);
  }
  ull bs = tid < ns ? base_of(j) : 0ull;
  for (;;) {
    // the next tile: further in this segment, or the next non-empty segment
    ull jn = j, posn = pos + (ull)m, enn = en;
    bool more = true;
    if (posn >= en) {
      jn = j + 1;
      ull s2 = 0, e2 = 0;
      while (jn < j1) {
        seg_bounds(O1, nblk, spb, nsb, jn, s2, e2);
        if (e2 > s2) break;
        jn++;
      }
      more = jn < j1;
      posn = s2;
      enn = e2;
    }
    Key kn[Q] = {};
    ull bn = 0;
    int mn = 0;
    if (more) {
      mn = (int)(enn - posn < (ull)TQ ? enn - posn : (ull)TQ);
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int idx = tid + q * NT;
        kn[q] = idx < mn ? Sin[posn + idx] : Key(0);
      }
      if (tid < ns && jn != j) bn = base_of(jn);
    }
    if (tid < ns) L.h[tid] = 0;
    const int np = m > tid ? (m - tid - 1) / NT + 1 : 0;
    __syncthreads();
    uint32_t rk[(Q + 1) / 2] = {};
#pragma unroll
    for (int q = 0; q < Q; q++)
      if (q < np) rk[q / 2] |= atomicAdd(&L.h[key_digit<Key>(kr[q], dpos, dbits)], 1u) << (16 * (q & 1));
    __syncthreads();
    const uint32_t v = tid < ns ? L.h[tid] : 0u;
    if (more && jn == j) bn = bs + v;
    const ull ex = block_excl_scan<NT>(v);
    if (tid < ns) {
      L.tcur[tid] = (uint32_t)ex;
      L.delta[tid] = (long long)bs - (long long)ex;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
      if (q < np) {
        const uint32_t d = key_digit<Key>(kr[q], dpos, dbits);
        L.T1[L.tcur[d] + ((rk[q / 2] >> (16 * (q & 1))) & 0xFFFFu)] = kr[q];
      }
    }
    __syncthreads();
    for (int p = tid; p < m; p += NT) {
      const Key w = L.T1[p];
      Sout[L.delta[key_digit<Key>(w, dpos, dbits)] + p] = w;
    }
    __syncthreads();
    if (!more) break;
#pragma unroll
    for (int q = 0; q < Q; q++) kr[q] = kn[q];
    bs = bn;
    j = jn;
    pos = posn;
    en = enn;
    m = mn;
  }
}

// O2T[((a * nsb + g) << db) + s] = O2[((a << db) + s) * nsb + g]: per
// level-1 bucket a, the (digit, segment) matrix of bases transposed through
// LDS in 32 x 32 tiles (grid: segment tiles, digit tiles, a).
__global__ __launch_bounds__(256) void k_seg_transpose(const ull* O2, uint64_t nsb, int db, ull* O2T) {
  __shared__ ull t[32][33];
  const uint64_t a = blockIdx.z;
  const int S = 1 << db;
  const uint64_t g0 = (uint64_t)blockIdx.x * 32;
  const int s0 = (int)blockIdx.y * 32;
  const int tx = (int)threadIdx.x & 31, ty = (int)threadIdx.x >> 5;
  const ull* src = O2 + ((a << db) * nsb);
  for (int r = ty; r < 32; r += 8) {
    const int sd = s0 + r;
    const uint64_t g = g0 + (uint64_t)tx;
    if (sd < S && g < nsb) t[r][tx] = src[(uint64_t)sd * nsb + g];
  }
  __syncthreads();
  ull* dst = O2T + ((a * nsb) << db);
  for (int r = ty; r < 32; r += 8) {
    const uint64_t g = g0 + (uint64_t)r;
    const int sd = s0 + tx;
    if (sd < S && g < nsb) dst[(g << db) + (uint64_t)sd] = t[tx][r];
  }
}

// boff2[(a << db) | s] = the first segment's base, O2T[((a * nsb) << db) + s];
// boff2[end] = total
__global__ void k_seg_boff(const ull* O2T, uint64_t nsb, int db, ull n2, ull total, ull* boff2) {
  const ull g = (ull)blockIdx.x * blockDim.x + threadIdx.x;
  const ull mask = (1ull << db) - 1;
  if (g < n2) boff2[g] = O2T[(((g >> db) * nsb) << db) + (g & mask)];
  if (g == n2) boff2[n2] = total;
}

// ---------------------------------------------------------------------------
// Per-bucket count: LDS hash table of the bucket's distinct keys (64-bit CAS,
// linear probing; duplicates cost one atomic add), then a counting sort of
// the distinct keys by a 12-bit sub-digit into LDS, a per-sub-bin insertion
// sort (sub-bins hold ~1-2 distinct keys) and a coalesced write of the
// bucket's (key, count) pairs.  This is the reference's sort + run-length +
// merge + aggregate (kmer_counting.cpp:3057-3086, 254-352, 169-250) for the
// keys of one bucket.

constexpr int BS_NT = 256;
constexpr int SUB_MAX_BITS = 9;
constexpr int SUB_BINS = 1 << SUB_MAX_BITS;

template <class Key>
struct HashCfg;
template <>
struct HashCfg<uint64_t> {
  static constexpr int T = 2048;      // table slots
  static constexpr int LIMIT = 1536;  // distinct keys per task (load 0.75)
  static constexpr int SUB = 9;       // sub-digit bits of the ordering pass
  static constexpr int BLOCKS = 6;    // blocks per CU the LDS allows (register cap follows)
};
template <>
struct HashCfg<u128> {
  static constexpr int T = 1024;
  static constexpr int LIMIT = 768;
  static constexpr int SUB = 8;
  static constexpr int BLOCKS = 5;
};
// tasks with more instances than this are split without trying
constexpr uint64_t TASK_MAX_INSTANCES = 1u << 18;
// linear-probe bound of an insert: a table too full to place a key within
// it sends the task to the next pass split (the host sizes buckets for a
// load of about 0.75, where probes average a few slots)
constexpr int HASH_PROBES = 256;
// a probe chain longer than STOP_AFTER slots re-reads the task's overflow
// flag every 8 probes: once one thread has found the table full, the others
// stop probing it (short chains, the usual case, never read it)
constexpr int STOP_AFTER = 8;

template <class Key>
struct HashTable;

// u64 keys are stored as themselves; EMPTY is a key outside the task.
template <>
struct HashTable<uint64_t> {
  static constexpr int T = HashCfg<uint64_t>::T;
  ull* hk;
  uint32_t* hc;
  ull empty;
  __device__ __forceinline__ void init(int tid, int nt) {
    for (int i = tid; i < T; i += nt) {
      hk[i] = empty;
      hc[i] = 0;
    }
  }
  // returns false when no free slot is found within HASH_PROBES probes, or
  // when a long chain finds the table already reported full (*stop)
  // (a read before the CAS, as for u128 keys, was slower here: a new key
  // then waits on two LDS round trips)
  __device__ __forceinline__ bool try_slot(uint32_t slot, uint64_t k) {
    const ull prev = atomicCAS(&hk[slot], empty, (ull)k);
    if (prev == empty || prev == (ull)k) {
      atomicAdd(&hc[slot], 1u);
      return true;
    }
    return false;
  }
  __device__ __forceinline__ bool insert(uint64_t k, const volatile int* stop) {
    uint32_t slot = (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> 40) & (T - 1);
    for (int probe = 0; probe < STOP_AFTER; probe++) {
      if (try_slot(slot, k)) return true;
      slot = (slot + 1) & (T - 1);
    }
    for (int probe = STOP_AFTER; probe < HASH_PROBES; probe++) {
      if ((probe & 7) == 0 && *stop) return false;
      if (try_slot(slot, k)) return true;
      slot = (slot + 1) & (T - 1);
    }
    return false;
  }
  __device__ __forceinline__ bool used(int i) const { return hc[i] != 0u; }
  __device__ __forceinline__ uint64_t key(int i) const { return hk[i]; }
};

// u128 keys: claim by 64-bit fingerprint, full key stored beside it and
// verified after all inserts (a fingerprint collision sends the task to the
// split path, which separates the two keys).
template <>
struct HashTable<u128> {
  static constexpr int T = HashCfg<u128>::T;
  static constexpr ull EMPTY = ~0ull;
  ull* hf;
  u128* hkf;
  uint32_t* hc;
  ull empty;
  // the high word through one odd multiply (a bijection) folded into the
  // low word, then one multiply-xorshift round for the top bits the home
  // slot takes: two 64-bit products instead of the four of two full mixes
  // (the table is bound by instruction issue).  Distinct keys of a task
  // collide with probability ~2^-64 per pair; a collision only sends the
  // task to the split path.
  __device__ __forceinline__ static ull fp(u128 k) {
    ull f = (ull)k ^ ((ull)(k >> 64) * 0x9e3779b97f4a7c15ull);
    f ^= f >> 29;
    f *= 0xbf58476d1ce4e5b9ull;
    return f ^ (f >> 32);
  }
  __device__ __forceinline__ void init(int tid, int nt) {
    for (int i = tid; i < T; i += nt) {
      hf[i] = EMPTY;
      hc[i] = 0;
    }
  }
  // home slot from the fingerprint's top bits, scaled to T (T need not be a
  // power of two); probes step to the next slot, wrapping
  __device__ __forceinline__ uint32_t home(ull f) const { return (uint32_t)(((f >> 32) * (ull)T) >> 32); }
  __device__ __forceinline__ static uint32_t next(uint32_t slot) { return slot + 1 == (uint32_t)T ? 0u : slot + 1; }
  // a slot changes only once, from EMPTY to its fingerprint: a plain read
  // finds a repeated key or steps over another one without an atomic, and
  // only an empty slot is claimed by CAS (k = 64 bacterial: 2.73 -> 2.56 ms)
  __device__ __forceinline__ bool try_slot(uint32_t slot, ull f, u128 k) {
    ull prev = __hip_atomic_load(&hf[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (prev == EMPTY) {
      prev = atomicCAS(&hf[slot], EMPTY, f);
      if (prev == EMPTY) hkf[slot] = k;
    }
    if (prev == EMPTY || prev == f) {
      atomicAdd(&hc[slot], 1u);
      return true;
    }
    return false;
  }
  // the slot that counts k, or -1 (table full, stopped, or a fingerprint
  // equal to EMPTY)
  __device__ __forceinline__ int insert_at(u128 k, const volatile int* stop) {
    const ull f = fp(k);
    if (f == EMPTY) return -1;
    uint32_t slot = home(f);
    for (int probe = 0; probe < STOP_AFTER; probe++) {
      if (try_slot(slot, f, k)) return (int)slot;
      slot = next(slot);
    }
    for (int probe = STOP_AFTER; probe < HASH_PROBES; probe++) {
      if ((probe & 7) == 0 && *stop) return -1;
      if (try_slot(slot, f, k)) return (int)slot;
      slot = next(slot);
    }
    return -1;
  }
  __device__ __forceinline__ bool insert(u128 k, const volatile int* stop) { return insert_at(k, stop) >= 0; }
  // the first probe of insert_at alone: the home slot if it holds f or is
  // claimed for it, else -1 (the slot holds another fingerprint)
  __device__ __forceinline__ int try_home(u128 k, ull f) {
    const uint32_t slot = home(f);
    return try_slot(slot, f, k) ? (int)slot : -1;
  }
  // the rest of insert_at, for a key whose home slot holds another fingerprint
  __device__ __forceinline__ int insert_after_home(u128 k, ull f, const volatile int* stop) {
    uint32_t slot = next(home(f));
    for (int probe = 1; probe < STOP_AFTER; probe++) {
      if (try_slot(slot, f, k)) return (int)slot;
      slot = next(slot);
    }
    for (int probe = STOP_AFTER; probe < HASH_PROBES; probe++) {
      if ((probe & 7) == 0 && *stop) return -1;
      if (try_slot(slot, f, k)) return (int)slot;
      slot = next(slot);
    }
    return -1;
  }
  // after a barrier: does key k sit in its fingerprint's slot?
  __device__ __forceinline__ bool verify(u128 k) const {
    const ull f = fp(k);
    uint32_t slot = home(f);
    for (int probe = 0; probe < T; probe++) {
      if (hf[slot] == f) return hkf[slot] == k;
      if (hf[slot] == EMPTY) return false;
      slot = next(slot);
    }
    return false;
  }
  __device__ __forceinline__ bool used(int i) const { return hc[i] != 0u; }
  __device__ __forceinline__ u128 key(int i) const { return hkf[i]; }
};

template <class Key>
struct HashLds;
template <>
struct HashLds<uint64_t> {
  ull hk[HashCfg<uint64_t>::T];
  uint32_t hc[HashCfg<uint64_t>::T];
  __device__ __forceinline__ HashTable<uint64_t> table(ull empty) {
    HashTable<uint64_t> t;
    t.hk = hk;
    t.hc = hc;
    t.empty = empty;
    return t;
  }
  __device__ __forceinline__ uint64_t* out_keys() { return reinterpret_cast<uint64_t*>(hk); }
};
template <>
struct HashLds<u128> {
  ull hf[HashCfg<u128>::T];
  u128 hkf[HashCfg<u128>::T];
  uint32_t hc[HashCfg<u128>::T];
  __device__ __forceinline__ HashTable<u128> table(ull) {
    HashTable<u128> t;
    t.hf = hf;
    t.hkf = hkf;
    t.hc = hc;
    t.empty = 0;
    return t;
  }
  __device__ __forceinline__ u128* out_keys() { return hkf; }
};

// Tasks: [tstart[i], tstart[i] + tcount[i]) of S, whose keys share their top
// `fixed` bits (fixed >= 1).  Writes the task's unique (key, count) pairs in
// ascending order to R / C at the task's start and tu[i] = their number.  A
// task with more distinct keys than a table holds is processed in 2^s passes,
// pass p taking the keys whose next s bits equal p (s = 1, 2, ... up to
// MAX_PASS_BITS); beyond that, or on a u128 fingerprint collision, tu[i] = -1
// and the task index goes to ovf_list for the host-driven split.
constexpr int MAX_PASS_BITS = 4;

template <class Key>
__global__ __launch_bounds__(BS_NT, HashCfg<Key>::BLOCKS) void k_bucket_hash(const Key* S, Key* R, int32_t* C,
                                                       const ull* tstart, const ull* tcount,
                                                       uint64_t ntasks, int K, int fixed,
                                                       long long* tu, ull* ovf_count, ull* ovf_list) {
  constexpr int T = HashCfg<Key>::T;
  constexpr int PER = T / BS_NT;  // table slots per thread
  constexpr int HB = sizeof(Key) == 8 ? 16 : 8;
  __shared__ HashLds<Key> lds;
  __shared__ uint32_t hs[SUB_BINS];
  __shared__ int over;
  __shared__ int qn;  // u128: keys of the batch queued for their probe chains

  const uint64_t task = blockIdx.x;
  if (task >= ntasks) return;
  const uint64_t start = tstart[task];
  const uint64_t n = tcount[task];
  if (n == 0) {
    if (threadIdx.x == 0) tu[task] = 0;
    return;
  }
  const int rem = 2 * K - fixed;
  // EMPTY for u64 tables: the task's first key with its top key bit flipped
  // (all keys of the task share that bit, so this value is not in the task)
  const Key k0 = S[start];
  const ull empty = (ull)(k0 ^ (Key(1) << (2 * K - 1)));
  HashTable<Key> ht = lds.table(empty);
  const int max_s = rem < MAX_PASS_BITS ? rem : MAX_PASS_BITS;
  bool done = false;
  ull out = 0;
  // a first try with all keys is pointless when there are too many of them
  int s = 0;
  while (s < max_s && (n >> s) > (uint64_t)TASK_MAX_INSTANCES) s++;
  for (; s <= max_s && !done; s++) {
    out = 0;
    bool failed = false;
    for (int pass = 0; pass < (1 << s) && !failed; pass++) {
      ht.init(threadIdx.x, BS_NT);
      if (threadIdx.x == 0) over = 0;
      __syncthreads();
      // keys are loaded HB at a time into registers so that the global-load
      // latency is paid once per batch, not once per insert; a task of one
      // batch keeps them there for the fingerprint check below
      Key kb[HB];
      int ksl[HB];  // u128: the slot each key of the (last) batch landed in
      int ov = 0;   // u128: over as read after the last batch's final barrier
      for (uint64_t b = 0; b < n; b += (uint64_t)BS_NT * HB) {
#pragma unroll
        for (int j = 0; j < HB; j++) {
          const uint64_t idx = b + threadIdx.x + (uint64_t)j * BS_NT;
          kb[j] = idx < n ? S[start + idx] : Key(0);
        }
        if constexpr (sizeof(Key) == 16) {
          // A wave runs each probe step as long as its slowest lane needs
          // it, so a batch goes in two phases: every key tries its home slot
          // (one probe, the common case), and the keys whose home slot holds
          // another fingerprint queue up (in hs, idle until the output
          // phase) and are then probed by all lanes together, so a long
          // chain holds up one queue round instead of each of the eight
          // keys of its lane.  Every thread takes the overflow exit on the
          // value read after the previous batch's last barrier (nothing
          // writes over between that read and the next barrier), so all
          // leave the loop together.
          if (ov) break;
          if (threadIdx.x == 0) qn = 0;
          __syncthreads();
          int qp[HB];
#pragma unroll
          for (int j = 0; j < HB; j++) {
            qp[j] = -1;
            const uint64_t idx = b + threadIdx.x + (uint64_t)j * BS_NT;
            if (idx < n && (int)key_digit<Key>(kb[j], rem - s, s) == pass) {
              const ull f = HashTable<u128>::fp(kb[j]);
              ksl[j] = f == HashTable<u128>::EMPTY ? -2 : ht.try_home(kb[j], f);
              if (ksl[j] == -2) {
                over = 1;
              } else if (ksl[j] < 0) {
                const int q = atomicAdd(&qn, 1);
                if (q < SUB_BINS) {
                  hs[q] = (uint32_t)(threadIdx.x + j * BS_NT);
                  qp[j] = q;
                } else {
                  ksl[j] = ht.insert_after_home(kb[j], f, &over);
                  if (ksl[j] < 0) over = 1;
                }
              }
            }
          }
          __syncthreads();
          const int nq = qn < SUB_BINS ? qn : SUB_BINS;
          for (int q = threadIdx.x; q < nq; q += BS_NT) {
            const u128 k = S[start + b + hs[q]];
            const int sl = ht.insert_after_home(k, HashTable<u128>::fp(k), &over);
            if (sl < 0) over = 1;
            hs[q] = (uint32_t)sl;
          }
          __syncthreads();
          ov = over;
#pragma unroll
          for (int j = 0; j < HB; j++)
            if (qp[j] >= 0) ksl[j] = (int)hs[qp[j]];
        } else {
          if (*(volatile int*)&over) break;
#pragma unroll
          for (int j = 0; j < HB; j++) {
            const uint64_t idx = b + threadIdx.x + (uint64_t)j * BS_NT;
            if (idx < n && (int)key_digit<Key>(kb[j], rem - s, s) == pass && !ht.insert(kb[j], &over)) over = 1;
          }
        }
      }
      __syncthreads();
      if constexpr (sizeof(Key) == 16) {
        if (!over) {
          // a key counted in a slot whose stored key differs shares its
          // fingerprint with that key: the task goes to the split path.  A
          // one-batch task checks the slots its inserts returned; a longer
          // one reloads its keys and finds their slots again.
          const bool one_batch = n <= (uint64_t)BS_NT * HB;
          for (uint64_t b = 0; b < n; b += (uint64_t)BS_NT * HB) {
            if (!one_batch) {
#pragma unroll
              for (int j = 0; j < HB; j++) {
                const uint64_t idx = b + threadIdx.x + (uint64_t)j * BS_NT;
                kb[j] = idx < n ? S[start + idx] : Key(0);
              }
            }
#pragma unroll
            for (int j = 0; j < HB; j++) {
              const uint64_t idx = b + threadIdx.x + (uint64_t)j * BS_NT;
              if (idx < n && (int)key_digit<Key>(kb[j], rem - s, s) == pass &&
                  !(one_batch ? ht.key(ksl[j]) == kb[j] : ht.verify(kb[j])))
                over = 2;
            }
          }
        }
        __syncthreads();
      }
      if (over) {
        // too many distinct keys: retry with one more split bit; a
        // fingerprint collision (over == 2) goes to the host split
        failed = true;
        if (over == 2) s = max_s;
        __syncthreads();
        break;
      }
      // order the distinct keys: counting sort by the sub-digit below the
      // pass bits, then rank inside each sub-bin (1-2 keys on average)
      const int r2 = rem - s;
      const int dbits = r2 < SUB_MAX_BITS ? r2 : SUB_MAX_BITS;
      const int dpos = r2 - dbits;
      const int nbins = 1 << dbits;
      Key rk[PER];
      uint32_t rc[PER], ri[PER];  // ri: index of the key inside its sub-bin
      ull ndk;  // distinct keys of the pass
      for (int i = threadIdx.x; i < nbins; i += BS_NT) hs[i] = 0;
      __syncthreads();
#pragma unroll
      for (int j = 0; j < PER; j++) {
        const int slot = threadIdx.x * PER + j;
        rc[j] = lds.hc[slot];
        rk[j] = rc[j] ? ht.key(slot) : Key(0);
        ri[j] = rc[j] ? atomicAdd(&hs[key_digit<Key>(rk[j], dpos, dbits)], 1u) : 0u;
      }
      __syncthreads();
      {
        constexpr int BPT = SUB_BINS / BS_NT;
        uint32_t v[BPT];
        ull sm = 0;
#pragma unroll
        for (int i = 0; i < BPT; i++) {
          const int bb = threadIdx.x * BPT + i;
          v[i] = bb < nbins ? hs[bb] : 0u;
          sm += v[i];
        }
        ull ex = block_scan_total<BS_NT>(sm, &ndk);
#pragma unroll
        for (int i = 0; i < BPT; i++) {
          const int bb = threadIdx.x * BPT + i;
          if (bb < nbins) {
            hs[bb] = (uint32_t)ex;
            ex += v[i];
          }
        }
      }
      __syncthreads();
      Key* ok = lds.out_keys();
      uint32_t* oc = lds.hc;
#pragma unroll
      for (int j = 0; j < PER; j++) {
        if (rc[j]) {
          const uint32_t pos = hs[key_digit<Key>(rk[j], dpos, dbits)] + ri[j];
          ok[pos] = rk[j];
          oc[pos] = rc[j];
        }
      }
      __syncthreads();
      const uint32_t nd = (uint32_t)ndk;
      for (uint32_t i = threadIdx.x; i < nd; i += BS_NT) {
        const Key v = ok[i];
        const int d = (int)key_digit<Key>(v, dpos, dbits);
        // the sub-bin's slots: from its start to the next one's (no end
        // array, so that six tables fit a CU's LDS)
        const int a = (int)hs[d];
        const int e = d + 1 < nbins ? (int)hs[d + 1] : (int)nd;
        // the sub-bin's first four slots read at once (most hold one or two
        // keys), the rest of a longer sub-bin one by one.  a < nd <= T, so a
        // read past the key array lands in the counts that follow it in the
        // same LDS struct, and only slots below e count.
        const Key w0 = ok[a], w1 = ok[a + 1], w2 = ok[a + 2], w3 = ok[a + 3];
        int rank = (int)(w0 < v) + (int)(a + 1 < e && w1 < v) + (int)(a + 2 < e && w2 < v) +
                   (int)(a + 3 < e && w3 < v);
        for (int j = a + 4; j < e; j++) rank += ok[j] < v;
        R[start + out + a + rank] = v;
        C[start + out + a + rank] = (int32_t)oc[i];
      }
      out += nd;
      __syncthreads();
    }
    if (!failed) done = true;
  }
  if (threadIdx.x == 0) {
    if (done) {
      tu[task] = (long long)out;
    } else {
      tu[task] = -1;
      ovf_list[atomicAdd(ovf_count, 1ull)] = task;
    }
  }
}

// All keys of the task are identical (no bits left to split on).
template <class Key>
__global__ void k_single_key(const Key* S, Key* R, int32_t* C, const ull* tstart, const ull* tcount,
                             uint64_t ntasks, long long* tu) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntasks) return;
  if (tcount[i] == 0) {
    tu[i] = 0;
    return;
  }
  R[tstart[i]] = S[tstart[i]];
  C[tstart[i]] = (int32_t)(uint32_t)tcount[i];
  tu[i] = 1;
}

// Split oversize tasks by the pbits-wide digit at bit pos (host-driven).
constexpr int PT_NT = 256;
constexpr int PT_PER = 4096;

template <class Key>
__global__ __launch_bounds__(PT_NT) void k_part_hist(const Key* S, const ull* tstart,
                                                     const ull* tcount, const uint32_t* blk_task,
                                                     const ull* blk_off, int pos, int pbits,
                                                     ull* th) {
  __shared__ uint32_t h[256];
  const int nb = 1 << pbits;
  for (int i = threadIdx.x; i < nb; i += PT_NT) h[i] = 0;
  __syncthreads();
  const uint32_t t = blk_task[blockIdx.x];
  const uint64_t s = tstart[t], n = tcount[t];
  const uint64_t off = blk_off[blockIdx.x];
  for (uint64_t i = off + threadIdx.x; i < n && i < off + PT_PER; i += PT_NT)
    atomicAdd(&h[key_digit<Key>(S[s + i], pos, pbits)], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += PT_NT)
    if (h[i]) atomicAdd(&th[(uint64_t)t * 256 + i], (ull)h[i]);
}

template <class Key>
__global__ __launch_bounds__(PT_NT) void k_part_scatter(const Key* S, const ull* tstart,
                                                        const ull* tcount, const uint32_t* blk_task,
                                                        const ull* blk_off, int pos, int pbits,
                                                        ull* tcur, Key* T) {
  __shared__ uint32_t h[256];
  __shared__ ull base[256];
  const int nb = 1 << pbits;
  for (int i = threadIdx.x; i < nb; i += PT_NT) h[i] = 0;
  __syncthreads();
  const uint32_t t = blk_task[blockIdx.x];
  const uint64_t s = tstart[t], n = tcount[t];
  const uint64_t off = blk_off[blockIdx.x];
  for (uint64_t i = off + threadIdx.x; i < n && i < off + PT_PER; i += PT_NT)
    atomicAdd(&h[key_digit<Key>(S[s + i], pos, pbits)], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += PT_NT)
    base[i] = h[i] ? atomicAdd(&tcur[(uint64_t)t * 256 + i], (ull)h[i]) : 0ull;
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += PT_NT) h[i] = 0;
  __syncthreads();
  for (uint64_t i = off + threadIdx.x; i < n && i < off + PT_PER; i += PT_NT) {
    const Key k = S[s + i];
    const uint32_t d = key_digit<Key>(k, pos, pbits);
    const uint32_t r = atomicAdd(&h[d], 1u);
    T[s + base[d] + r] = k;
  }
}

template <class Key>
__global__ void k_copy_tasks(Key* S, const Key* T, const ull* tstart, const ull* tcount,
                             const uint32_t* blk_task, const ull* blk_off) {
  const uint32_t t = blk_task[blockIdx.x];
  const uint64_t s = tstart[t], n = tcount[t];
  const uint64_t off = blk_off[blockIdx.x];
  for (uint64_t i = off + threadIdx.x; i < n && i < off + PT_PER; i += blockDim.x) S[s + i] = T[s + i];
}

// Dense table from per-leaf results (leaves in key order).
template <class Key>
__global__ void k_compact(const Key* S, const int32_t* C, const ull* lstart, const long long* lu,
                          const ull* loff, uint64_t nleaves, Key* TK, int32_t* TC) {
  const uint64_t b = blockIdx.x;
  if (b >= nleaves) return;
  const long long u = lu[b];
  const uint64_t s = lstart[b], o = loff[b];
  for (long long i = threadIdx.x; i < u; i += blockDim.x) {
    TK[o + i] = S[s + i];
    TC[o + i] = C[s + i];
  }
}

// ---------------------------------------------------------------------------
// Cutoff (kmer_histogram.cpp:84-140).

constexpr int CUT_NT = 1024;
constexpr int T_MAX = 4096;
// A cutoff block covers CUT_R rounds of CUT_NT * 4 counts: in round r thread
// t reads the four at r * CUT_NT * 4 + 4 t (one 16-byte load, coalesced),
// and its loads of all rounds are in flight together.
constexpr int CUT_R = 4;
constexpr uint64_t CUT_PER = (uint64_t)CUT_R * CUT_NT * 4;

__device__ __forceinline__ void cut_load4(const int32_t* TC, uint64_t U, uint64_t base, int32_t fill,
                                          int32_t* v) {
  if (base + 4 <= U) {
    const int4 q = *reinterpret_cast<const int4*>(TC + base);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = base + i < U ? TC[base + i] : fill;
  }
}

__global__ __launch_bounds__(CUT_NT) void k_count_hist(const int32_t* TC, uint64_t U, int t, ull* hist) {
  __shared__ uint32_t h[T_MAX];
  for (int i = threadIdx.x; i < t; i += CUT_NT) h[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * CUT_PER;
  for (uint64_t i = base + threadIdx.x; i < U && i < base + CUT_PER; i += CUT_NT) {
    const int v = TC[i];
    if (v >= 0 && v < t) atomicAdd(&h[v], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < t; i += CUT_NT)
    if (h[i]) atomicAdd(&hist[i], (ull)h[i]);
}

// Small-t path (t <= CUT_SMALL_T): per-wave LDS sub-histograms (no
// contention between waves), the block's histogram stored to BH[b * t + c]
// without global atomics; column sums give hist, and the kept count of block
// b is its size minus its entries with count < min_bucket.
constexpr int CUT_SMALL_T = 64;

__global__ __launch_bounds__(CUT_NT) void k_count_hist_blocks(const int32_t* TC, uint64_t U, int t,
                                                             uint32_t* BH) {
  constexpr int NW = CUT_NT / 64;
  __shared__ uint32_t h[NW * CUT_SMALL_T];
  const int w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < NW * t; i += CUT_NT) h[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * CUT_PER + (uint64_t)threadIdx.x * 4;
  int32_t v[CUT_R * 4];
#pragma unroll
  for (int r = 0; r < CUT_R; r++) cut_load4(TC, U, base + (uint64_t)r * CUT_NT * 4, -1, v + 4 * r);
  // count 1 (the k-mers seen once, most of the table) is summed per wave by
  // ballot instead of 64 lanes hitting one LDS word
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < CUT_R * 4; i++) {
    const bool one = v[i] == 1 && t > 1;
    const unsigned long long m = __ballot(one);
    if (lane == 0 && m) atomicAdd(&h[w * t + 1], (uint32_t)__popcll(m));
    if (!one && v[i] >= 0 && v[i] < t) atomicAdd(&h[w * t + v[i]], 1u);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < t; c += CUT_NT) {
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) sum += h[q * t + c];
    BH[(uint64_t)blockIdx.x * t + c] = sum;
  }
}

// hist[c] = sum over blocks of BH[b * t + c]; one block per column
__global__ __launch_bounds__(1024) void k_hist_columns(const uint32_t* BH, uint64_t nb, int t, ull* hist) {
  const int c = blockIdx.x;
  ull s = 0;
  for (uint64_t b = threadIdx.x; b < nb; b += 1024) s += BH[b * t + c];
  s = block_sum<1024>(s);
  if (threadIdx.x == 0) hist[c] = s;
}

// kept[b] = block size - entries with count < min_bucket
__global__ void k_kept_from_hist(const uint32_t* BH, uint64_t nb, uint64_t U, int t, const int* mb,
                                 ull* kept) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int m = *mb;
  const uint64_t size = (b + 1) * CUT_PER <= U ? CUT_PER : U - b * CUT_PER;
  uint64_t below = 0;
  for (int c = 0; c < m && c < t; c++) below += BH[b * t + c];
  kept[b] = size - below;
}

// min_bucket = first argmin over hist[1..t-1]; an empty range gives t.
__global__ void k_min_bucket(const ull* hist, int t, int* mb) {
  if (threadIdx.x || blockIdx.x) return;
  int best = t;
  if (t > 1) {
    best = 1;
    for (int i = 2; i < t; i++)
      if (hist[i] < hist[best]) best = i;
  }
  *mb = best;
}

__global__ __launch_bounds__(CUT_NT) void k_kept_count(const int32_t* TC, uint64_t U, const int* mb,
                                                      ull* bsum) {
  const int thr = *mb;
  const uint64_t base = (uint64_t)blockIdx.x * CUT_PER + (uint64_t)threadIdx.x * 4;
  int32_t v[CUT_R * 4];
#pragma unroll
  for (int r = 0; r < CUT_R; r++) cut_load4(TC, U, base + (uint64_t)r * CUT_NT * 4, INT32_MIN, v + 4 * r);
  ull c = 0;
#pragma unroll
  for (int i = 0; i < CUT_R * 4; i++) c += v[i] >= thr;
  c = block_sum<CUT_NT>(c);
  if (threadIdx.x == 0) bsum[blockIdx.x] = c;
}

template <class Key>
__global__ __launch_bounds__(CUT_NT) void k_kept_write(const Key* TK, const int32_t* TC, uint64_t U,
                                                      const int* mb, const ull* boffs, Key* KK,
                                                      int32_t* KC) {
  // a block with no kept entry reads nothing (high cutoffs keep a few
  // thousand of a billion k-mers); the test is the same for every thread
  if (boffs[blockIdx.x + 1] == boffs[blockIdx.x]) return;
  const int thr = *mb;
  const uint64_t base = (uint64_t)blockIdx.x * CUT_PER + (uint64_t)threadIdx.x * 4;
  int32_t cv[CUT_R * 4];
#pragma unroll
  for (int r = 0; r < CUT_R; r++) cut_load4(TC, U, base + (uint64_t)r * CUT_NT * 4, INT32_MIN, cv + 4 * r);
  // the block's kept entries go out in index order: round by round, and in
  // a round by thread.  A thread's four round counts (<= 4 each, <= 4096
  // per round over the block) travel through one scan in 16-bit fields.
  ull packed = 0;
#pragma unroll
  for (int r = 0; r < CUT_R; r++) {
    ull c = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) c += cv[4 * r + i] >= thr;
    packed |= c << (16 * r);
  }
  ull tot;
  const ull ex = block_scan_total<CUT_NT>(packed, &tot);
  ull rbase = boffs[blockIdx.x];
#pragma unroll
  for (int r = 0; r < CUT_R; r++) {
    ull o = rbase + ((ex >> (16 * r)) & 0xFFFFull);
    const uint64_t b = base + (uint64_t)r * CUT_NT * 4;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (cv[4 * r + i] >= thr) {
        KK[o] = TK[b + i];
        KC[o] = cv[4 * r + i];
        o++;
      }
    }
    rbase += (tot >> (16 * r)) & 0xFFFFull;
  }
}

// ---------------------------------------------------------------------------
// Macro nodes (mn_construct.cpp:212-477).
//
// The kept table KK[0..V) is ascending.  pred(x) = x >> 2 is non-decreasing
// over it; within quarter a (first base a, indices [q[a], q[a+1])) succ(x) =
// x & SUCC_MASK is ascending.  Node keys are the union of these five sorted
// lists.  Range g covers node keys [s_g, s_{g+1}) with s_g = pred(KK[g*PC]).

constexpr int MN_NT = 512;
// 768 pred entries per range keep a u64 block at 36 KB of LDS, four blocks
// of 512 threads per CU (1024 / 3072 allowed three: 0.88 -> 0.82 ms for the
// bacterial nodes)
constexpr int MN_PC = 768;     // pred entries per range
constexpr int MN_ECAP = 2304;  // LDS capacity (entries of all five lists)

struct MnPlan {
  ull lo[5];
  ull hi[5];
  ull base;
  ull e;
};

template <class Key>
__device__ __forceinline__ Key mn_pred(Key x) { return x >> 2; }

template <class Key>
__global__ void k_mn_quarters(const Key* KK, uint64_t V, int K, ull* q) {
  if (threadIdx.x >= 5 || blockIdx.x) return;
  const int a = threadIdx.x;
  if (a == 0) { q[0] = 0; return; }
  if (a == 4) { q[4] = V; return; }
  const Key lim = Key(a) << (2 * K - 2);
  uint64_t lo = 0, hi = V;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (KK[mid] < lim) lo = mid + 1; else hi = mid;
  }
  q[a] = lo;
}

template <class Key>
__device__ __forceinline__ uint64_t lb_pred(const Key* KK, uint64_t lo, uint64_t hi, Key x) {
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (mn_pred(KK[mid]) < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
template <class Key>
__device__ __forceinline__ uint64_t lb_succ(const Key* KK, uint64_t lo, uint64_t hi, Key x, Key m) {
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if ((KK[mid] & m) < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <class Key>
__global__ void k_mn_plan(const Key* KK, uint64_t V, int K, const ull* q, uint64_t G, MnPlan* plan,
                          ull* maxe) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const Key msk = key_mask<Key>(2 * (K - 1));
  const bool first = g == 0, last = g + 1 == G;
  const Key s_lo = first ? Key(0) : mn_pred(KK[g * MN_PC]);
  const Key s_hi = last ? Key(0) : mn_pred(KK[(g + 1) * MN_PC]);
  MnPlan p;
  p.lo[0] = first ? 0 : lb_pred(KK, 0, V, s_lo);
  p.hi[0] = last ? V : lb_pred(KK, 0, V, s_hi);
  ull base = p.lo[0], e = p.hi[0] - p.lo[0];
  for (int a = 0; a < 4; a++) {
    p.lo[a + 1] = first ? q[a] : lb_succ(KK, q[a], q[a + 1], s_lo, msk);
    p.hi[a + 1] = last ? q[a + 1] : lb_succ(KK, q[a], q[a + 1], s_hi, msk);
    base += p.lo[a + 1] - q[a];
    e += p.hi[a + 1] - p.lo[a + 1];
  }
  p.base = base;
  p.e = e;
  plan[g] = p;
  atomicMax(maxe, e);
}

// Node record as written to memory (matches pk_mnode).
struct MnRec {
  ull key_lo, key_hi;
  int32_t pre[4];
  int32_t suf[4];
};

template <class Key>
__device__ __forceinline__ void set_key(MnRec& r, Key k);
template <>
__device__ __forceinline__ void set_key<uint64_t>(MnRec& r, uint64_t k) {
  r.key_lo = k;
  r.key_hi = 0;
}
template <>
__device__ __forceinline__ void set_key<u128>(MnRec& r, u128 k) {
  r.key_lo = (ull)k;
  r.key_hi = (ull)(k >> 64);
}

// List view: 5 sorted lists of node keys, either staged in LDS (lk) or
// read from KK through pred/succ.  List 0 = pred keys (may repeat, up to 4
// equal), lists 1..4 = succ keys of quarter a = l - 1 (no repeats).
template <class Key, bool GLOBAL>
struct MnLists {
  const Key* KK;
  Key msk;
  const Key* lk;  // LDS copy (GLOBAL == false)
  ull off[5];     // list start (index into lk, or into KK)
  ull len[5];
  __device__ __forceinline__ Key at(ull off_l, bool pred, ull i) const {
    if (GLOBAL) {
      const Key x = KK[off_l + i];
      return pred ? mn_pred(x) : (x & msk);
    }
    return lk[off_l + i];
  }
};

template <class Key, bool GLOBAL>
__device__ __forceinline__ ull mn_lb(const MnLists<Key, GLOBAL>& L, ull o, ull n, bool pred, Key x) {
  ull lo = 0, hi = n;
  while (lo < hi) {
    const ull mid = (lo + hi) >> 1;
    if (L.at(o, pred, mid) < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
template <class Key, bool GLOBAL>
__device__ __forceinline__ ull mn_ub(const MnLists<Key, GLOBAL>& L, ull o, ull n, bool pred, Key x) {
  ull lo = 0, hi = n;
  while (lo < hi) {
    const ull mid = (lo + hi) >> 1;
    if (L.at(o, pred, mid) <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Positions of x in all five lists, walked forward from the previous
// element's positions (x never decreases along a thread's chunk of one list).
template <class Key, bool GLOBAL>
struct MnCursor {
  ull lb[5], ub[5];
  __device__ __forceinline__ void seek(const MnLists<Key, GLOBAL>& L, Key x) {
#pragma unroll
    for (int o = 0; o < 5; o++) {
      lb[o] = mn_lb<Key, GLOBAL>(L, L.off[o], L.len[o], o == 0, x);
      ub[o] = lb[o];
      while (ub[o] < L.len[o] && L.at(L.off[o], o == 0, ub[o]) <= x) ub[o]++;
    }
  }
  __device__ __forceinline__ void advance(const MnLists<Key, GLOBAL>& L, Key x) {
#pragma unroll
    for (int o = 0; o < 5; o++) {
      while (lb[o] < L.len[o] && L.at(L.off[o], o == 0, lb[o]) < x) lb[o]++;
      if (ub[o] < lb[o]) ub[o] = lb[o];
      while (ub[o] < L.len[o] && L.at(L.off[o], o == 0, ub[o]) <= x) ub[o]++;
    }
  }
};

// LDS ranges (GLOBAL == false) use the launch's dynamic LDS: the range's
// five lists (node keys) | per rank: start bit 31, list << 16 | index |
// per node: rank of its first element.
template <class Key>
constexpr size_t mn_lds_bytes() {
  return MN_ECAP * (sizeof(Key) + 2 * sizeof(uint32_t));
}

// One block per range.  GLOBAL == false: ranges with e <= MN_ECAP, lists in
// LDS.  GLOBAL == true: the remaining ranges, lists and rank flags in global
// memory (gflags, indexed from plan.base).  Each thread handles a contiguous
// chunk of the concatenated lists and walks its positions in the other lists
// (pass 1: rank of every element in the merged order and whether it starts
// a node); pass 2 numbers the nodes; pass 3 writes the records.
template <class Key, bool GLOBAL>
__global__ __launch_bounds__(MN_NT) void k_mn_build(const Key* KK, const int32_t* KC, int K,
                                                    const MnPlan* plan, uint64_t G, MnRec* NT,
                                                    ull* ng, uint32_t* gflags) {
  extern __shared__ __align__(16) unsigned char mn_smem[];
  Key* lk = reinterpret_cast<Key*>(mn_smem);
  uint32_t* lflag = reinterpret_cast<uint32_t*>(mn_smem + MN_ECAP * sizeof(Key));
  uint32_t* lstart = lflag + MN_ECAP;
  __shared__ ull chunk_base;
  __shared__ ull s_lo[5];   // list starts in the kept table (block-uniform,
  __shared__ ull s_cum[6];  // indexed at run time from LDS, not registers)
  const uint64_t g = blockIdx.x;
  if (g >= G) return;
  const MnPlan p = plan[g];
  const bool big = p.e > (ull)MN_ECAP;
  if (big != GLOBAL) return;
  const ull E = p.e;
  MnLists<Key, GLOBAL> L;
  L.KK = KK;
  L.msk = key_mask<Key>(2 * (K - 1));
  L.lk = lk;
  ull cum[6];
  {
    ull o = 0;
#pragma unroll
    for (int l = 0; l < 5; l++) {
      L.len[l] = p.hi[l] - p.lo[l];
      L.off[l] = GLOBAL ? p.lo[l] : o;
      cum[l] = o;
      o += L.len[l];
    }
    cum[5] = o;
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      s_lo[l] = p.lo[l];
      s_cum[l] = cum[l];
    }
    s_cum[5] = cum[5];
  }
  if (!GLOBAL) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      for (ull i = threadIdx.x; i < L.len[l]; i += MN_NT) {
        const Key x = KK[p.lo[l] + i];
        lk[cum[l] + i] = l == 0 ? mn_pred(x) : (x & L.msk);
      }
    }
  }
  __syncthreads();
  uint32_t* flags = GLOBAL ? gflags + p.base : lflag;
  const ull chunk = (E + MN_NT - 1) / MN_NT;
  const ull e0 = (ull)threadIdx.x * chunk;
  const ull e1 = e0 + chunk < E ? e0 + chunk : E;
  // pass 1: merged rank and node-start flag of every element of the chunk
#pragma unroll
  for (int l = 0; l < 5; l++) {
    const ull a = e0 > cum[l] ? e0 : cum[l];
    const ull b = e1 < cum[l + 1] ? e1 : cum[l + 1];
    if (a >= b) continue;
    MnCursor<Key, GLOBAL> cur;
    for (ull e = a; e < b; e++) {
      const ull j = e - cum[l];
      const Key x = L.at(L.off[l], l == 0, j);
      if (e == a) cur.seek(L, x); else cur.advance(L, x);
      ull rank = j;
      bool start = cur.lb[l] == j;
#pragma unroll
      for (int o = 0; o < 5; o++) {
        if (o < l) {
          rank += cur.ub[o];
          if (cur.ub[o] != cur.lb[o]) start = false;
        } else if (o > l) {
          rank += cur.lb[o];
        }
      }
      if (GLOBAL) flags[rank] = start ? 1u : 0u;
      else flags[rank] = (start ? 0x80000000u : 0u) | ((uint32_t)l << 16) | (uint32_t)j;
    }
  }
  __syncthreads();
  // pass 2: exclusive scan of the start flags in rank order (chunks of
  // MN_NT*4).  Global ranges keep "node index + 1" per start rank; LDS
  // ranges keep their per-rank element ids and note each node's first rank.
  if (threadIdx.x == 0) chunk_base = 0;
  __syncthreads();
  for (ull c0 = 0; c0 < E; c0 += (ull)MN_NT * 4) {
    const ull i0 = c0 + (ull)threadIdx.x * 4;
    uint32_t v[4];
    ull sm = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      v[i] = i0 + i < E ? (GLOBAL ? flags[i0 + i] : flags[i0 + i] >> 31) : 0u;
      sm += v[i];
    }
    ull ex = block_excl_scan<MN_NT>(sm) + chunk_base;
    const ull tot = block_sum<MN_NT>(sm);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (GLOBAL) {
        if (i0 + i < E) flags[i0 + i] = v[i] ? (uint32_t)(ex + 1) : 0u;
      } else if (v[i]) {
        lstart[ex] = (uint32_t)(i0 + i);
      }
      ex += v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) chunk_base += tot;
    __syncthreads();
  }
  const ull nodes = chunk_base;
  if (!GLOBAL) {
    // pass 3 (LDS ranges): one thread per node, in node order, so that
    // consecutive lanes write consecutive records.  A node's elements are
    // consecutive in rank order: its first, then the ranks up to the next
    // start.  Suffix b of the node: pred-list element with last base b;
    // prefix a: the element of quarter list a + 1.
    for (ull nd = threadIdx.x; nd < nodes; nd += MN_NT) {
      const uint32_t r0 = lstart[nd];
      const uint32_t id0 = lflag[r0];
      MnRec rec;
      set_key<Key>(rec, lk[s_cum[(id0 >> 16) & 7u] + (id0 & 0xFFFFu)]);
      int32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0, q0 = 0, q1 = 0, q2 = 0, q3 = 0;
      for (ull r = r0; r < E; r++) {
        const uint32_t id = lflag[r];
        if (r != r0 && (id >> 31)) break;
        const uint32_t l = (id >> 16) & 7u;
        const ull gi = s_lo[l] + (id & 0xFFFFu);
        const int32_t cnt = KC[gi];
        if (l == 0) {
          const uint32_t bse = (uint32_t)(KK[gi] & Key(3));
          s0 = bse == 0 ? cnt : s0;
          s1 = bse == 1 ? cnt : s1;
          s2 = bse == 2 ? cnt : s2;
          s3 = bse == 3 ? cnt : s3;
        } else {
          q0 = l == 1 ? cnt : q0;
          q1 = l == 2 ? cnt : q1;
          q2 = l == 3 ? cnt : q2;
          q3 = l == 4 ? cnt : q3;
        }
      }
      rec.suf[0] = s0;
      rec.suf[1] = s1;
      rec.suf[2] = s2;
      rec.suf[3] = s3;
      rec.pre[0] = q0;
      rec.pre[1] = q1;
      rec.pre[2] = q2;
      rec.pre[3] = q3;
      NT[p.base + nd] = rec;
    }
  } else {
    // pass 3 (global ranges): every node start assembles its record
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const ull a = e0 > cum[l] ? e0 : cum[l];
      const ull b = e1 < cum[l + 1] ? e1 : cum[l + 1];
      if (a >= b) continue;
      MnCursor<Key, GLOBAL> cur;
      for (ull e = a; e < b; e++) {
        const ull j = e - cum[l];
        const Key x = L.at(L.off[l], l == 0, j);
        if (e == a) cur.seek(L, x); else cur.advance(L, x);
        if (cur.lb[l] != j) continue;
        ull rank = j;
        bool start = true;
#pragma unroll
        for (int o = 0; o < 5; o++) {
          if (o < l) {
            rank += cur.ub[o];
            if (cur.ub[o] != cur.lb[o]) start = false;
          } else if (o > l) {
            rank += cur.lb[o];
          }
        }
        if (!start) continue;
        const uint32_t idx1 = flags[rank];
        MnRec rec;
        set_key<Key>(rec, x);
        int32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (ull i = cur.lb[0]; i < cur.ub[0]; i++) {
          const ull gi = p.lo[0] + i;
          const uint32_t bse = (uint32_t)(KK[gi] & Key(3));
          const int32_t cnt = KC[gi];
          s0 = bse == 0 ? cnt : s0;
          s1 = bse == 1 ? cnt : s1;
          s2 = bse == 2 ? cnt : s2;
          s3 = bse == 3 ? cnt : s3;
        }
        rec.suf[0] = s0;
        rec.suf[1] = s1;
        rec.suf[2] = s2;
        rec.suf[3] = s3;
#pragma unroll
        for (int q = 0; q < 4; q++)
          rec.pre[q] = cur.ub[q + 1] > cur.lb[q + 1] ? KC[p.lo[q + 1] + cur.lb[q + 1]] : 0;
        NT[p.base + idx1 - 1] = rec;
      }
    }
  }
  if (threadIdx.x == 0) ng[g] = nodes;
}

// Initial wiring of one macro node per thread (MacroNode::setup_wiring,
// includes/distribute_kmers.h:550-656, after initiate_mnode_wiring appends
// the terminals, src/wire.cpp:84-96).  Entry lists: prefixes in the given
// base order, suffixes ascending, each followed by its terminal; counts are
// (count, visit) pairs.  The index sorts (Comp_rev: visit, then count,
// descending) run on at most five entries, where libstdc++'s std::sort is an
// insertion sort, i.e. stable -- so is the insertion sort here.
__device__ __forceinline__ void wire_sort(int* idx, int n, const int (*cnt)[2]) {
  for (int i = 1; i < n; i++) {
    const int v = idx[i];
    int j = i;
    while (j > 0) {
      const int u = idx[j - 1];
      const bool before = cnt[v][1] > cnt[u][1] || (cnt[v][1] == cnt[u][1] && cnt[v][0] > cnt[u][0]);
      if (!before) break;
      idx[j] = u;
      j--;
    }
    idx[j] = v;
  }
}

__global__ void k_mn_wire(const MnRec* nodes, uint64_t M, const uint8_t* pre_order, int cov,
                          pk_mnwire* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const MnRec r = nodes[i];
  int pc_[6][2], sc_[6][2];
  int np = 0, ns = 0;
  for (int q = 0; q < 4; q++) {
    int b = q;
    if (pre_order) {
      const uint8_t o = pre_order[i * 4 + q];
      if (o > 3) continue;
      b = o;
    }
    if (r.pre[b]) {
      pc_[np][0] = r.pre[b];
      pc_[np][1] = (r.pre[b] + cov - 1) / cov;
      np++;
    }
  }
  for (int b = 0; b < 4; b++)
    if (r.suf[b]) {
      sc_[ns][0] = r.suf[b];
      sc_[ns][1] = (r.suf[b] + cov - 1) / cov;
      ns++;
    }
  int pc = 0, sc = 0;
  for (int a = 0; a < np; a++) pc += pc_[a][1];
  for (int a = 0; a < ns; a++) sc += sc_[a][1];
  // terminals (count -1 in the reference until setup_wiring balances them)
  pc_[np][0] = 1;
  pc_[np][1] = sc - pc > 0 ? sc - pc : 0;
  sc_[ns][0] = 1;
  sc_[ns][1] = pc - sc > 0 ? pc - sc : 0;
  np++;
  ns++;
  int ip[6], is[6];
  for (int a = 0; a < np; a++) ip[a] = a;
  for (int a = 0; a < ns; a++) is[a] = a;
  wire_sort(ip, np, pc_);
  wire_sort(is, ns, sc_);
  pk_mnwire w;
  w.n_prefix = np;
  w.n_suffix = ns;
  w.pre_terminal_count[0] = pc_[np - 1][0];
  w.pre_terminal_count[1] = pc_[np - 1][1];
  w.suf_terminal_count[0] = sc_[ns - 1][0];
  w.suf_terminal_count[1] = sc_[ns - 1][1];
  for (int a = 0; a < 11; a++) w.wire[a] = pk_wire_info{0, 0, 0};
  for (int a = 0; a < 5; a++) w.prefix_begin[a][0] = w.prefix_begin[a][1] = 0;
  int off[6] = {0, 0, 0, 0, 0, 0};
  int leftover = sc + sc_[ns - 1][1];
  int top_p = 0, top_s = 0, var_p = 0, var_s = 0, wi = 0, p_size = 0, last_pid = -1, begin = -1;
  while (leftover > 0 && wi < 11 && top_p < np && top_s < ns) {
    const int sid = is[top_s], pid = ip[top_p];
    const int a = pc_[pid][1] - var_p, b = sc_[sid][1] - var_s;
    const int c = a < b ? a : b;
    w.wire[wi] = pk_wire_info{sid, off[sid], c};
    if (last_pid != pid) {
      begin = wi;
      last_pid = pid;
    }
    p_size++;
    wi++;
    leftover -= c;
    var_p += c;
    var_s += c;
    off[sid] += c;
    if (var_p == pc_[pid][1]) {
      var_p = 0;
      top_p++;
      if (pid < 5) {
        w.prefix_begin[pid][0] = begin;
        w.prefix_begin[pid][1] = p_size;
      }
      p_size = 0;
    }
    if (var_s == sc_[sid][1]) {
      var_s = 0;
      top_s++;
    }
  }
  w.n_wire = wi;
  out[i] = w;
}

__global__ void k_mn_compact(const MnRec* NT, const MnPlan* plan, const ull* ng, const ull* noff,
                             uint64_t G, MnRec* out) {
  const uint64_t g = blockIdx.x;
  if (g >= G) return;
  const ull b = plan[g].base, o = noff[g], c = ng[g];
  // a record is three 16-byte words (both arrays start 256-byte aligned):
  // copy the range as words, one per lane
  static_assert(sizeof(MnRec) == 48, "MnRec is three 16-byte words");
  const uint4* src = reinterpret_cast<const uint4*>(NT + b);
  uint4* dst = reinterpret_cast<uint4*>(out + o);
  for (ull i = threadIdx.x; i < 3 * c; i += blockDim.x) dst[i] = src[i];
}

// Macro-node share of rank `rank` at P ranks: nodes whose (k-1)-mer key is
// owned by the rank (retrieve_proc_id(pred / succ), mn_construct.cpp:
// 234-247; the rank receives every kept k-mer of its nodes,
// kmer_histogram.cpp:140-170, so its records are the one-rank records).
template <class Key>
__global__ void k_node_share_flags(const MnRec* nodes, uint64_t M, int P, int rank, uint32_t* flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  Key key;
  if (sizeof(Key) == 8) key = (Key)nodes[i].key_lo;
  else key = (Key)(((u128)nodes[i].key_hi << 64) | (u128)nodes[i].key_lo);
  flag[i] = key_owner(key, P) == rank ? 1u : 0u;
}

__global__ void k_node_share_write(const MnRec* nodes, uint64_t M, const uint32_t* flag, const ull* pos,
                                   MnRec* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M || !flag[i]) return;
  out[pos[i]] = nodes[i];
}

// ---------------------------------------------------------------------------
// Host side.

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

enum BufId {
  B_FASTA, B_BLK, B_BLKOFF, B_NL, B_RSTART, B_RLEN, B_KC, B_KOFS, B_SMALL, B_BCNT, B_BOFF, B_BCUR,
  B_S, B_C, B_TU, B_UOFF, B_TK, B_TC, B_HIST, B_KBS, B_KBOFF, B_KK, B_KC2, B_MNQ, B_MNPLAN,
  B_MNNT, B_MNNG, B_MNNOFF, B_MNFLAG, B_NODES, B_TMP, B_OVF, B_PT0, B_PT1, B_PT2, B_PT3, B_PT4,
  B_PT5, B_PT6, B_BLKGT, B_S2, B_BCNT2, B_BOFF2, B_BCUR2, B_TP, B_HIST1, B_OFF1, B_HIST2, B_OFF2, B_CUTBH, B_LFREQ, B_WORDER, B_WIRE, B_NORM, B_NORMNL, B_NORMLEN, B_NORMOFF, B_NORMLS, B_NORMLL, B_SEGH, B_SEGO, B_SEGT, B_SCAN0, B_SCAN1 = B_SCAN0 + 8, B_COUNT = B_SCAN0 + 16
};

}  // namespace

struct pk_ctx {
  int device = 0;
  int ncu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint64_t held = 0;
  std::string err;
  DevBuf buf[B_COUNT];
  // 0-19: stage boundaries; 20 + 4 L ..: further partition level L (hist
  // start/end, scatter start/end) for the first MS_TIMED_LEVELS levels
  hipEvent_t ev[20 + 4 * 4] = {};
  int ms_levels = 0;  // further partition levels of the last count
  int ms_form[MS_TIMED_LEVELS] = {};  // their scatter forms (timed levels)
  pk_timing timing{};
  // state of the current table
  int k = 0;
  int words = 1;
  uint64_t U = 0, V = 0, M = 0, N = 0, R = 0;
  bool have_table = false, have_kept = false, have_nodes = false;
  int t = 0, mb = 0;
  int key_bits = 0;  // bits of a table key: 2k (main path) or the word (builder)
  // stages whose event times have not been read yet
  bool pend_count = false, pend_split = false, pend_cutoff = false, pend_mn = false;
  // pinned staging of host input: PIN_THREADS host threads, two buffers each
  void* pin[8] = {};
  hipEvent_t pin_ev[8] = {};
  bool pin_used[8] = {};
  bool pend_norm = false;  // normalize_ms of the last pk_count_format
  uint64_t norm_n = 0;     // bytes of the last rewritten input (B_NORM)
  bool have_norm = false;
  pk_tuning tuning{};  // zeros: built-in choices
  // Scratch mode of pk_count_noalloc: need() carves buffers from `arena`
  // (the device's static scratch array) instead of allocating; the caller's
  // output arrays stand in for the table (ids marked in arena_fixed).
  bool arena_mode = false;
  char* arena = nullptr;
  uint64_t arena_cap = 0, arena_used = 0;
  bool arena_fixed[B_COUNT] = {};
  uint8_t win_lut[256] = {};  // byte -> code of the last pk_builder_count_windows
  // buffers from the device's stream-ordered pool (pk_ctx_set_alloc_mode)
  bool stream_alloc = false;

  int fail(int code, const std::string& what) {
    err = what;
    return code;
  }
  int hip_fail(hipError_t e, const char* what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return PK_EHIP;
  }
};

#define PK_HIP(ctx, x)                                   \
  do {                                                   \
    hipError_t e_ = (x);                                 \
    if (e_ != hipSuccess) return (ctx)->hip_fail(e_, #x); \
  } while (0)

#define PK_TRY(x)          \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

namespace {

template <class T>
int need(pk_ctx* c, int id, uint64_t count, T** out) {
  DevBuf& b = c->buf[id];
  const uint64_t bytes = count * sizeof(T) + 256;
  if (c->arena_mode) {
    if (c->arena_fixed[id]) {
      // a caller's output array: it must hold count entries (an empty
      // table asks for one entry and writes none)
      if (count * sizeof(T) > b.cap && !(count == 1 && c->U == 0))
        return c->fail(PK_ENOMEM, "output arrays hold fewer entries than the table (" + std::to_string(count) +
                                      " needed)");
    } else if (b.cap < bytes) {
      const uint64_t take = (bytes + 255) & ~255ull;
      if (c->arena_used + take > c->arena_cap)
        return c->fail(PK_ENOMEM, "input needs more than the " + std::to_string(c->arena_cap >> 20) +
                                      " MiB scratch of pk_count_noalloc");
      b.p = c->arena + c->arena_used;
      b.cap = take;
      c->arena_used += take;
    }
    *out = reinterpret_cast<T*>(b.p);
    return PK_OK;
  }
  if (b.cap < bytes) {
    if (b.p) {
      // stream-ordered contexts free on their own stream: no device-wide wait
      hipError_t e = c->stream_alloc ? hipFreeAsync(b.p, c->stream) : hipFree(b.p);
      (void)e;
      c->held -= b.cap;
      b.p = nullptr;
      b.cap = 0;
    }
    if (c->budget && c->held + bytes > c->budget)
      return c->fail(PK_ENOMEM, "device budget exceeded");
    void* p = nullptr;
    hipError_t e = c->stream_alloc ? hipMallocAsync(&p, bytes, c->stream) : hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return c->fail(PK_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    b.p = p;
    b.cap = bytes;
    c->held += bytes;
  }
  *out = reinterpret_cast<T*>(b.p);
  return PK_OK;
}

inline unsigned grid_for(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

int fill_u64(pk_ctx* c, ull* p, ull v, uint64_t n) {
  if (!n) return PK_OK;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, p, v, n);
  PK_HIP(c, hipGetLastError());
  return PK_OK;
}

// out[0..n) = exclusive scan of in, out[n] = total.  level selects scratch.
template <class T>
int scan_excl(pk_ctx* c, const T* in, ull* out, uint64_t n, int level = 0) {
  if (n == 0) {
    return fill_u64(c, out, 0, 1);
  }
  if (level >= 8) return c->fail(PK_EINVAL, "scan too deep");
  const uint64_t nb = (n + SCAN_PER - 1) / SCAN_PER;
  if (nb == 1) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_scan_apply<T>), dim3(1), dim3(SCAN_NT), 0, c->stream, in,
                       out, (const ull*)nullptr, n);
    PK_HIP(c, hipGetLastError());
    return PK_OK;
  }
  ull* sums;
  PK_TRY(need(c, B_SCAN0 + level, nb + 1, &sums));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_scan_reduce<T>), dim3((unsigned)nb), dim3(SCAN_NT), 0,
                     c->stream, in, sums, n);
  PK_HIP(c, hipGetLastError());
  PK_TRY(scan_excl<ull>(c, sums, sums, nb, level + 1));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_scan_apply<T>), dim3((unsigned)nb), dim3(SCAN_NT), 0,
                     c->stream, in, out, (const ull*)sums, n);
  PK_HIP(c, hipGetLastError());
  // total = sums[nb]
  hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(1), 0, c->stream, out + n, sums + nb);
  PK_HIP(c, hipGetLastError());
  return PK_OK;
}

int d2h(pk_ctx* c, void* dst, const void* src, size_t bytes) {
  PK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  PK_HIP(c, hipStreamSynchronize(c->stream));
  return PK_OK;
}

// Bucket bits: at least 1 (the hash table's EMPTY value flips the top key
// bit), and enough that a bucket holds about 3 tables of instances (a bucket
// with more distinct keys than a table holds takes extra passes).
int choose_bits(uint64_t N, int K, int cap_keys, float factor) {
  int b = 1;
  const uint64_t target = (uint64_t)((double)cap_keys * factor);
  while (b < 24 && b < 2 * K && (N >> b) > target) b++;
  return b;
}


struct Leaf {
  ull start;
  long long u;
};

// Recursively split oversize tasks; returns the leaves (in key order) for
// each input task.
template <class Key>
int resolve_overflow(pk_ctx* c, Key* S, Key* R, int32_t* C, int K, int fixed_bits,
                     const std::vector<ull>& starts, const std::vector<ull>& counts,
                     std::vector<std::vector<Leaf>>& leaves) {
  constexpr int CAP = HashCfg<Key>::LIMIT;
  leaves.assign(starts.size(), {});
  const int rem = 2 * K - fixed_bits;
  if (rem <= 0) {
    // all keys of every task are identical
    const uint64_t nt = starts.size();
    ull *ts, *tc;
    long long* tu;
    PK_TRY(need(c, B_PT0, nt, &ts));
    PK_TRY(need(c, B_PT1, nt, &tc));
    PK_TRY(need(c, B_PT2, nt, &tu));
    PK_HIP(c, hipMemcpyAsync(ts, starts.data(), nt * 8, hipMemcpyHostToDevice, c->stream));
    PK_HIP(c, hipMemcpyAsync(tc, counts.data(), nt * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_single_key<Key>), dim3(grid_for(nt, 256)), dim3(256), 0,
                       c->stream, S, R, C, ts, tc, nt, tu);
    PK_HIP(c, hipGetLastError());
    std::vector<long long> hu(nt);
    PK_TRY(d2h(c, hu.data(), tu, nt * 8));
    for (size_t i = 0; i < nt; i++) leaves[i].push_back(Leaf{starts[i], hu[i]});
    return PK_OK;
  }
  // digit width for this level
  ull maxn = 0;
  for (auto n : counts) maxn = std::max<ull>(maxn, n);
  int pbits = 1;
  while (pbits < 8 && (maxn >> pbits) > (ull)(CAP / 2)) pbits++;
  if (pbits > rem) pbits = rem;
  const int pos = rem - pbits;
  // block map
  std::vector<uint32_t> bt;
  std::vector<ull> bo;
  for (size_t t = 0; t < starts.size(); t++)
    for (ull o = 0; o < counts[t]; o += PT_PER) {
      bt.push_back((uint32_t)t);
      bo.push_back(o);
    }
  const uint64_t nt = starts.size(), nblk = bt.size();
  ull *ts, *tc, *bo_d, *th;
  uint32_t* bt_d;
  Key* T;
  PK_TRY(need(c, B_PT0, nt, &ts));
  PK_TRY(need(c, B_PT1, nt, &tc));
  PK_TRY(need(c, B_PT3, nblk, &bt_d));
  PK_TRY(need(c, B_PT4, nblk, &bo_d));
  PK_TRY(need(c, B_PT5, nt * 256, &th));
  PK_TRY(need(c, B_TMP, c->N, &T));
  PK_HIP(c, hipMemcpyAsync(ts, starts.data(), nt * 8, hipMemcpyHostToDevice, c->stream));
  PK_HIP(c, hipMemcpyAsync(tc, counts.data(), nt * 8, hipMemcpyHostToDevice, c->stream));
  PK_HIP(c, hipMemcpyAsync(bt_d, bt.data(), nblk * 4, hipMemcpyHostToDevice, c->stream));
  PK_HIP(c, hipMemcpyAsync(bo_d, bo.data(), nblk * 8, hipMemcpyHostToDevice, c->stream));
  PK_TRY(fill_u64(c, th, 0, nt * 256));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_part_hist<Key>), dim3((unsigned)nblk), dim3(PT_NT), 0,
                     c->stream, S, ts, tc, bt_d, bo_d, pos, pbits, th);
  PK_HIP(c, hipGetLastError());
  std::vector<ull> hh(nt * 256);
  PK_TRY(d2h(c, hh.data(), th, nt * 256 * 8));
  std::vector<ull> cur(nt * 256, 0);
  std::vector<ull> sub_start, sub_count;
  std::vector<std::pair<size_t, int>> owner;  // (task, digit)
  for (size_t t = 0; t < nt; t++) {
    ull o = 0;
    for (int d = 0; d < (1 << pbits); d++) {
      cur[t * 256 + d] = o;
      if (hh[t * 256 + d]) {
        sub_start.push_back(starts[t] + o);
        sub_count.push_back(hh[t * 256 + d]);
        owner.push_back({t, d});
      }
      o += hh[t * 256 + d];
    }
  }
  PK_HIP(c, hipMemcpyAsync(th, cur.data(), nt * 256 * 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_part_scatter<Key>), dim3((unsigned)nblk), dim3(PT_NT), 0,
                     c->stream, S, ts, tc, bt_d, bo_d, pos, pbits, th, T);
  PK_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_copy_tasks<Key>), dim3((unsigned)nblk), dim3(PT_NT), 0,
                     c->stream, S, T, ts, tc, bt_d, bo_d);
  PK_HIP(c, hipGetLastError());
  // sort the sub-tasks that fit
  const int fixed2 = fixed_bits + pbits;
  const uint64_t ns = sub_start.size();
  ull *ss, *sc, *ovc, *ovl;
  long long* su;
  PK_TRY(need(c, B_PT6, ns * 3 + 2, &ss));
  sc = ss + ns;
  su = reinterpret_cast<long long*>(ss + 2 * ns);
  PK_TRY(need(c, B_OVF, ns + 1, &ovl));
  PK_TRY(need(c, B_PT2, 1, &ovc));
  PK_HIP(c, hipMemcpyAsync(ss, sub_start.data(), ns * 8, hipMemcpyHostToDevice, c->stream));
  PK_HIP(c, hipMemcpyAsync(sc, sub_count.data(), ns * 8, hipMemcpyHostToDevice, c->stream));
  PK_TRY(fill_u64(c, ovc, 0, 1));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bucket_hash<Key>), dim3((unsigned)ns), dim3(BS_NT), 0,
                     c->stream, S, R, C, ss, sc, ns, K, fixed2, su, ovc, ovl);
  PK_HIP(c, hipGetLastError());
  std::vector<long long> hu(ns);
  PK_TRY(d2h(c, hu.data(), su, ns * 8));
  std::vector<ull> o_start, o_count;
  std::vector<size_t> o_idx;
  for (size_t i = 0; i < ns; i++)
    if (hu[i] < 0) {
      o_start.push_back(sub_start[i]);
      o_count.push_back(sub_count[i]);
      o_idx.push_back(i);
    }
  std::vector<std::vector<Leaf>> sub_leaves(ns);
  if (!o_start.empty()) {
    std::vector<std::vector<Leaf>> rec;
    PK_TRY(resolve_overflow<Key>(c, S, R, C, K, fixed2, o_start, o_count, rec));
    for (size_t j = 0; j < o_idx.size(); j++) sub_leaves[o_idx[j]] = rec[j];
  }
  for (size_t i = 0; i < ns; i++) {
    auto& dst = leaves[owner[i].first];
    if (hu[i] >= 0) dst.push_back(Leaf{sub_start[i], hu[i]});
    else dst.insert(dst.end(), sub_leaves[i].begin(), sub_leaves[i].end());
  }
  return PK_OK;
}

// Index of the event bytes ('\n', or '\n' and '>') of fa: *m events at
// *ev_out; *p0_out = the position of the first '>' (n when there is none).
// small[0] is scratch for the pass that looks for it in every block.
int event_index(pk_ctx* c, const uint8_t* fa, uint64_t n, uint8_t e0, uint8_t e1, ull* small,
                ull** ev_out, uint64_t* m_out, uint64_t* p0_out, int ev_buf = B_NL, int follow = 0) {
  hipStream_t st = c->stream;
  const uint64_t nb = (n + NL_PER_BLOCK - 1) / NL_PER_BLOCK;
  ull *blk, *blkoff, *nl, *blkgt;
  PK_TRY(need(c, B_BLK, nb, &blk));
  PK_TRY(need(c, B_BLKGT, nb + 1, &blkgt));
  PK_TRY(need(c, B_BLKOFF, nb + 1, &blkoff));
  if (nb) {
    hipLaunchKernelGGL(k_nl_count, dim3((unsigned)nb), dim3(NL_NT), 0, st, fa, n, e0, e1, blk, blkgt, 0);
    PK_HIP(c, hipGetLastError());
  }
  PK_TRY(scan_excl<ull>(c, blk, blkoff, nb));
  // a '>' in the first block is the input's first one (other blocks report
  // one only from their byte-wise tail, which comes after it)
  ull hm[2] = {0, ~0ull};
  PK_HIP(c, hipMemcpyAsync(&hm[0], blkoff + nb, 8, hipMemcpyDeviceToHost, st));
  if (nb) PK_HIP(c, hipMemcpyAsync(&hm[1], blkgt, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipStreamSynchronize(st));
  if (hm[1] == ~0ull && nb > 1) {
    // no '>' in the first block: the first one anywhere (same counts again)
    hipLaunchKernelGGL(k_nl_count, dim3((unsigned)nb), dim3(NL_NT), 0, st, fa, n, e0, e1, blk, blkgt, 1);
    PK_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(1024), 0, st, blkgt, nb, small);
    PK_HIP(c, hipGetLastError());
    PK_HIP(c, hipMemcpyAsync(&hm[1], small, 8, hipMemcpyDeviceToHost, st));
    PK_HIP(c, hipStreamSynchronize(st));
  }
  const uint64_t m = hm[0];
  PK_TRY(need(c, ev_buf, m + 1, &nl));
  if (nb) {
    hipLaunchKernelGGL(k_nl_write, dim3((unsigned)nb), dim3(NL_NT), 0, st, fa, n, e0, e1, blkoff, nl, follow);
    PK_HIP(c, hipGetLastError());
  }
  *ev_out = nl;
  *m_out = m;
  *p0_out = hm[1] == ~0ull ? n : hm[1];
  return PK_OK;
}

// The input rewritten into strict single-line FASTA (B_NORM, *out_n bytes).
int normalize_impl(pk_ctx* c, const uint8_t* in, uint64_t n, int fmt, const uint8_t** out_p, uint64_t* out_n) {
  hipStream_t st = c->stream;
  ull* small;
  PK_TRY(need(c, B_SMALL, 16, &small));
  // small[8] = first '>' (unused), [9] = first record line, [10] = first
  // malformed line, [11] = first over-long line
  PK_TRY(fill_u64(c, small + 8, ~0ull, 4));
  ull* nl;
  uint64_t m, p0;
  PK_TRY(event_index(c, in, n, '\n', '\n', small + 8, &nl, &m, &p0, B_NORMNL));
  // a last line without '\n' is a line; the empty tail after a final '\n' is not
  ull last_nl = 0;
  if (m) PK_TRY(d2h(c, &last_nl, nl + m - 1, 8));
  const uint64_t L = m + ((n > 0 && (m == 0 || last_nl + 1 < n)) ? 1 : 0);
  if (fmt == FMT_FASTQ && L % 4 != 0)
    return c->fail(PK_EFORMAT, "FASTQ input has " + std::to_string(L) + " lines, not a multiple of 4");
  uint32_t *olen, *llen;
  ull *ooff, *lstart;
  uint8_t* out;
  PK_TRY(need(c, B_NORMLEN, L + 1, &olen));
  PK_TRY(need(c, B_NORMOFF, L + 1, &ooff));
  PK_TRY(need(c, B_NORMLS, L + 1, &lstart));
  PK_TRY(need(c, B_NORMLL, L + 1, &llen));
  PK_TRY(need(c, B_NORM, n + L + 64, &out));
  if (L) {
    if (fmt == FMT_FASTA) {
      hipLaunchKernelGGL(k_fmt_first_record, dim3(grid_for(L, 256)), dim3(256), 0, st, in, n, nl, m, L, small + 9);
      PK_HIP(c, hipGetLastError());
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fmt_lines<FMT_FASTA>), dim3(grid_for(L, 256)), dim3(256), 0, st, in, n,
                         nl, m, L, small + 9, olen, lstart, llen, small + 10, small + 11);
    } else {
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fmt_lines<FMT_FASTQ>), dim3(grid_for(L, 256)), dim3(256), 0, st, in, n,
                         nl, m, L, small + 9, olen, lstart, llen, small + 10, small + 11);
    }
    PK_HIP(c, hipGetLastError());
  }
  PK_TRY(scan_excl<uint32_t>(c, olen, ooff, L));
  ull hv[3];
  PK_HIP(c, hipMemcpyAsync(&hv[0], ooff + L, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipMemcpyAsync(&hv[1], small + 10, 16, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipStreamSynchronize(st));
  if (hv[1] != ~0ull)
    return c->fail(PK_EFORMAT, "FASTQ record line " + std::to_string(hv[1]) +
                                   (hv[1] % 4 == 0 ? " does not start with '@'" : " does not start with '+'"));
  if (hv[2] != ~0ull) return c->fail(PK_EINVAL, "line " + std::to_string(hv[2]) + " longer than 4 GiB");
  if (L) {
    // '\n's before each index block: the scan event_index left in B_BLKOFF
    const ull* blk_nl = reinterpret_cast<const ull*>(c->buf[B_BLKOFF].p);
    const uint64_t nb = (n + NL_PER_BLOCK - 1) / NL_PER_BLOCK;
    if (fmt == FMT_FASTA)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fmt_copy<FMT_FASTA>), dim3((unsigned)nb), dim3(NL_NT), 0, st, in, n,
                         blk_nl, lstart, llen, olen, ooff, out);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fmt_copy<FMT_FASTQ>), dim3((unsigned)nb), dim3(NL_NT), 0, st, in, n,
                         blk_nl, lstart, llen, olen, ooff, out);
    PK_HIP(c, hipGetLastError());
  }
  *out_p = out;
  *out_n = hv[0];
  return PK_OK;
}

// Read table of the main path: records of the reference's Sliding_window.
int records_main(pk_ctx* c, const uint8_t* fa, uint64_t n, int K, ull** rs, uint32_t** rl,
                 uint64_t* R_out, uint64_t* N_out) {
  hipStream_t st = c->stream;
  ull* small;
  PK_TRY(need(c, B_SMALL, 16, &small));
  // small[0] = first '>', [1] = first unprocessed, [2] = first format error,
  // [3] = first over-long read, [4] = overflow count
  PK_TRY(fill_u64(c, small, ~0ull, 4));
  PK_TRY(fill_u64(c, small + 4, 0, 1));
  if (n >> NL_FLAG_SHIFT) return c->fail(PK_EINVAL, "input larger than 2^62 bytes");
  ull* nl;
  uint64_t m, p0;
  // entries carry what follows each '\n' (bits 62-63): only k_reads reads them
  PK_TRY(event_index(c, fa, n, '\n', '\n', small, &nl, &m, &p0, B_NL, 1));
  const uint64_t ncand = m / 2 + 2;
  ull *rstart, *kofs;
  uint32_t *rlen, *kc;
  PK_TRY(need(c, B_RSTART, ncand, &rstart));
  PK_TRY(need(c, B_RLEN, ncand, &rlen));
  PK_TRY(need(c, B_KC, ncand, &kc));
  PK_TRY(need(c, B_KOFS, ncand + 1, &kofs));
  hipLaunchKernelGGL(k_reads, dim3(grid_for(ncand, 256)), dim3(256), 0, st, n, nl, m, p0, K,
                     ncand, rstart, rlen, kc, small + 1, small + 2, small + 3);
  PK_HIP(c, hipGetLastError());
  PK_TRY(scan_excl<uint32_t>(c, kc, kofs, ncand));
  ull hs[4];
  PK_HIP(c, hipMemcpyAsync(hs, small + 1, 24, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipMemcpyAsync(&hs[3], kofs + ncand, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipStreamSynchronize(st));
  const uint64_t R = std::min<uint64_t>(hs[0], ncand);
  if (hs[1] != ~0ull && hs[1] <= R)
    return c->fail(PK_EFORMAT, "input is not strict single-line FASTA (record " +
                                   std::to_string(hs[1]) + " does not start with '>')");
  if (hs[2] != ~0ull && hs[2] < R) return c->fail(PK_EINVAL, "read longer than 4 GiB");
  *rs = rstart;
  *rl = rlen;
  *R_out = R;
  *N_out = hs[3];
  return PK_OK;
}

// Read table of the builder API (bld_record).
int records_builder(pk_ctx* c, const uint8_t* fa, uint64_t n, int klen, uint64_t read_len, BldSyms sy,
                    ull** rs, uint32_t** rl, uint64_t* R_out, uint64_t* N_out) {
  hipStream_t st = c->stream;
  ull* small;
  PK_TRY(need(c, B_SMALL, 16, &small));
  PK_TRY(fill_u64(c, small, ~0ull, 4));
  PK_TRY(fill_u64(c, small + 4, 0, 1));
  ull* ev;
  uint64_t m, p0;
  PK_TRY(event_index(c, fa, n, sy.delim, sy.skip, small, &ev, &m, &p0));
  ull *rstart, *kofs, *idx;
  uint32_t *rlen, *kc, *flag;
  PK_TRY(need(c, B_RSTART, m + 1, &rstart));
  PK_TRY(need(c, B_RLEN, m + 1, &rlen));
  PK_TRY(need(c, B_KC, m + 1, &kc));
  PK_TRY(need(c, B_KOFS, m + 2, &kofs));
  PK_TRY(need(c, B_TU, m + 1, &flag));
  PK_TRY(need(c, B_UOFF, m + 2, &idx));
  if (m) {
    hipLaunchKernelGGL(k_bld_flags, dim3(grid_for(m, 256)), dim3(256), 0, st, fa, n, ev, m, klen,
                       read_len, sy, flag);
    PK_HIP(c, hipGetLastError());
  }
  PK_TRY(scan_excl<uint32_t>(c, flag, idx, m));
  ull R = 0;
  PK_TRY(d2h(c, &R, idx + m, 8));
  if (m) {
    hipLaunchKernelGGL(k_bld_write, dim3(grid_for(m, 256)), dim3(256), 0, st, fa, n, ev, m, klen,
                       read_len, sy, idx, rstart, rlen, kc);
    PK_HIP(c, hipGetLastError());
  }
  PK_TRY(scan_excl<uint32_t>(c, kc, kofs, R));
  ull N = 0;
  PK_TRY(d2h(c, &N, kofs + R, 8));
  *rs = rstart;
  *rl = rlen;
  *R_out = R;
  *N_out = N;
  return PK_OK;
}

// Further partition levels from memory, the per-bucket count and the dense
// table.  S holds the N keys (KB key bits) split into B buckets by their top
// `fixed` bits, at offsets boff (the B_BOFF buffer); levels of <= MS_MAX_BITS
// follow until btot bits are fixed.  S2 is the other instance array, C the
// per-instance counts; small[4] (the overflow count) starts at zero.  set0 =
// 1 when boff is in B_BOFF2.
template <class Key>
int count_partitioned(pk_ctx* c, Key* S, Key* S2, int32_t* C, ull* boff, uint64_t B, int fixed, int btot,
                      int KB, uint64_t N, int set0 = 0) {
  Key* TK;
  int32_t* TC;
  hipStream_t st = c->stream;
  const pk_tuning& tun = c->tuning;
  ull* small;
  PK_TRY(need(c, B_SMALL, 16, &small));
  // set: which offset / histogram buffers the input level used (boff sits
  // in B_BOFF when 0, B_BOFF2 when 1)
  int set = set0;
  Key* Sin = S;
  Key* Sout = S2;
  constexpr int TQ = MS_NT * TileCfg<Key>::Q;
  while (fixed < btot) {
    // the remaining bits over as few levels of <= MS_MAX_BITS as possible,
    // evenly
    const int left = btot - fixed;
    const int nlev = (left + MS_MAX_BITS - 1) / MS_MAX_BITS;
    const int db = (left + nlev - 1) / nlev;
    const uint64_t B2 = B << db;
    const uint64_t tiles = N / TQ + B + 1;  // upper bound on tp[B]
    ull *boff2, *tp, *O2;
    uint32_t* H2;
    const int o = set ? 0 : 1;  // the other offset array
    PK_TRY(need(c, o ? B_BOFF2 : B_BOFF, B2 + 1, &boff2));
    PK_TRY(need(c, B_TP, B + 1, &tp));
    // alternate histogram buffers per level: the previous level's kernels
    // may still be reading theirs
    PK_TRY(need(c, o ? B_HIST2 : B_HIST1, tiles << db, &H2));
    PK_TRY(need(c, o ? B_OFF2 : B_OFF1, (tiles << db) + 1, &O2));
    const int lv = c->ms_levels < MS_TIMED_LEVELS ? c->ms_levels : -1;
    hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(1024), 0, st, boff, (int)B, TQ, tp);
    PK_HIP(c, hipGetLastError());
    if (lv >= 0) PK_HIP(c, hipEventRecord(c->ev[20 + 4 * lv], st));
    // grid and scan over the upper bound of the level's tiles: blocks past
    // tp[B] return at once, and the scan of their (unused) histogram slots
    // only extends past every offset that is read
    const ull ntiles = tiles;
    const int dpos = KB - fixed - db;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ms_hist<Key>), dim3((unsigned)ntiles), dim3(MS_NT), 0, st, Sin,
                       boff, tp, (int)B, dpos, db, H2);
    PK_HIP(c, hipGetLastError());
    if (lv >= 0) PK_HIP(c, hipEventRecord(c->ev[21 + 4 * lv], st));
    PK_TRY(scan_excl<uint32_t>(c, H2, O2, ntiles << db));
    hipLaunchKernelGGL(k_ms_boff, dim3(grid_for(B2 + 1, 256)), dim3(256), 0, st, O2, tp, (int)B, db,
                       (ull)N, boff2);
    PK_HIP(c, hipGetLastError());
    // a long level runs on persistent blocks, two per CU (a tile buffer
    // takes about 68 KB of LDS), once each gets >= 256 tiles: measured faster
    // at ~1400 tiles per block (k = 64, 2.9e9 instances), slower at ~60
    // (k = 32, 2.5e8 instances) than one block per tile
    const ull pgrid = 2ull * (ull)c->ncu;
    const bool persistent = tun.ms_scatter ? tun.ms_scatter == 1 : ntiles >= 256 * pgrid;
    if (lv >= 0) PK_HIP(c, hipEventRecord(c->ev[22 + 4 * lv], st));
    c->timing.ms_scatter_form = persistent ? 1 : 2;
    if (lv >= 0) c->ms_form[lv] = persistent ? 1 : 2;
    if (persistent)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ms_scatter_p<Key, MS_P_NT>), dim3((unsigned)pgrid), dim3(MS_P_NT), 0, st,
                         Sin, boff, tp, (int)B, dpos, db, O2, Sout);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ms_scatter<Key>), dim3((unsigned)ntiles), dim3(MS_NT), 0, st,
                         Sin, boff, tp, (int)B, dpos, db, O2, Sout);
    PK_HIP(c, hipGetLastError());
    if (lv >= 0) PK_HIP(c, hipEventRecord(c->ev[23 + 4 * lv], st));
    c->ms_levels++;
    std::swap(Sin, Sout);
    boff = boff2;
    B = B2;
    fixed += db;
    set = o;
  }
  S = Sin;
  Key* Res = Sout;  // the other array receives the (key, count) results
  // bucket sizes of the last level
  ull* bcnt;
  PK_TRY(need(c, B_BCNT, B, &bcnt));
  hipLaunchKernelGGL(k_diff, dim3(grid_for(B, 256)), dim3(256), 0, st, boff, B, bcnt);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[5], st));
  // 4. per-bucket count
  long long* tu;
  ull *ovl, *uoff;
  PK_TRY(need(c, B_TU, B, &tu));
  PK_TRY(need(c, B_OVF, B + 1, &ovl));
  PK_TRY(need(c, B_UOFF, B + 1, &uoff));
  PK_HIP(c, hipEventRecord(c->ev[6], st));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bucket_hash<Key>), dim3((unsigned)B), dim3(BS_NT), 0, st, S,
                     Res, C, boff, bcnt, B, KB / 2, btot, tu, small + 4, ovl);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[7], st));
  // the scan of the distinct counts goes ahead of the overflow check: with
  // no oversize bucket (the usual case) one synchronisation yields both
  PK_TRY(scan_excl<long long>(c, tu, uoff, B));
  ull hv[2];
  PK_HIP(c, hipMemcpyAsync(&hv[0], small + 4, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipMemcpyAsync(&hv[1], uoff + B, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipStreamSynchronize(st));
  const ull novf = hv[0];
  c->timing.n_split = (int32_t)novf;
  PK_HIP(c, hipEventRecord(c->ev[8], st));
  const ull* lstart = boff;
  const long long* lu = tu;
  uint64_t nleaves = B;
  if (novf) {
    // leaves = buckets, with each oversize bucket replaced by its split
    std::vector<ull> hbo(B + 1), hbc(B);
    std::vector<long long> htu(B);
    PK_TRY(d2h(c, hbo.data(), boff, (B + 1) * 8));
    PK_TRY(d2h(c, hbc.data(), bcnt, B * 8));
    PK_TRY(d2h(c, htu.data(), tu, B * 8));
    std::vector<ull> os, oc;
    std::vector<size_t> ob;
    for (uint64_t b = 0; b < B; b++)
      if (htu[b] < 0) {
        os.push_back(hbo[b]);
        oc.push_back(hbc[b]);
        ob.push_back(b);
      }
    std::vector<std::vector<Leaf>> rec;
    PK_TRY(resolve_overflow<Key>(c, S, Res, C, KB / 2, btot, os, oc, rec));
    std::vector<ull> ls;
    std::vector<long long> lus;
    size_t j = 0;
    for (uint64_t b = 0; b < B; b++) {
      if (j < ob.size() && ob[j] == b) {
        for (auto& lf : rec[j]) {
          ls.push_back(lf.start);
          lus.push_back(lf.u);
        }
        j++;
      } else {
        ls.push_back(hbo[b]);
        lus.push_back(htu[b]);
      }
    }
    nleaves = ls.size();
    ull* lsd;
    long long* lud;
    PK_TRY(need(c, B_PT0, nleaves, &lsd));
    PK_TRY(need(c, B_PT1, nleaves, &lud));
    PK_TRY(need(c, B_UOFF, nleaves + 1, &uoff));
    PK_HIP(c, hipMemcpyAsync(lsd, ls.data(), nleaves * 8, hipMemcpyHostToDevice, st));
    PK_HIP(c, hipMemcpyAsync(lud, lus.data(), nleaves * 8, hipMemcpyHostToDevice, st));
    lstart = lsd;
    lu = lud;
  }
  PK_HIP(c, hipEventRecord(c->ev[9], st));
  ull U = hv[1];
  if (novf) {
    PK_TRY(scan_excl<long long>(c, lu, uoff, nleaves));
    PK_TRY(d2h(c, &U, uoff + nleaves, 8));
  }
  c->U = U;
  PK_TRY(need(c, B_TK, U, &TK));
  PK_TRY(need(c, B_TC, U, &TC));
  PK_HIP(c, hipEventRecord(c->ev[10], st));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_compact<Key>), dim3((unsigned)nleaves), dim3(256), 0, st, Res,
                     C, lstart, lu, uoff, nleaves, TK, TC);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[11], st));
  // stage times are read from the events when asked for (pk_last_timing):
  // no wait here, the compaction runs on while the caller goes on
  c->pend_count = true;
  c->pend_split = novf != 0;
  c->have_table = true;
  return PK_OK;
}

// Count the k-mers of the read table (rstart, rlen)[0..R) (records in file
// order), N instances.
// K = window length, KB = key bits (2K on the main path, the whole word for
// the builder encoding), mask = word mask of the rolling update.
template <class Key, int ENC, bool OWN = false>
int count_records(pk_ctx* c, const uint8_t* fa, uint64_t n, const ull* rstart, const uint32_t* rlen,
                  uint64_t R, uint64_t N, int K, int KB, Key mask, OwnCfg cf = OwnCfg{nullptr, 1, 0, 0}) {
  constexpr int CAP = HashCfg<Key>::LIMIT;
  hipStream_t st = c->stream;
  ull* small;
  PK_TRY(need(c, B_SMALL, 16, &small));
  c->R = R;
  c->N = N;
  c->k = K;
  c->words = sizeof(Key) / 8;
  c->timing.n_instances = N;
  c->timing.n_reads = R;
  c->ms_levels = 0;
  PK_HIP(c, hipEventRecord(c->ev[1], st));
  Key *S, *TK;
  int32_t *C, *TC;
  if (N == 0) {
    c->U = 0;
    PK_TRY(need(c, B_TK, 1, &TK));
    PK_TRY(need(c, B_TC, 1, &TC));
    c->have_table = true;
    return PK_OK;
  }
  // 3. partition: level 1 fused with the parse, further levels from memory,
  // at most 8 bits per level, until a bucket holds ~1.5 hash tables' worth
  // of instances.  Write bases come from scanned digit-major histograms.
  // with owner routing, about N / P instances are this rank's
  const pk_tuning& tun = c->tuning;
  const float bf = tun.bucket_factor > 0.f ? tun.bucket_factor : 3.0f;
  const int btot = choose_bits(OWN ? (N + cf.P - 1) / cf.P : N, KB / 2, CAP, bf);
  int d1;
  if (tun.l1_bits > 0 && tun.l1_bits <= L1_MAX_BITS) {
    d1 = btot < tun.l1_bits ? btot : tun.l1_bits;
  } else if (btot > L1_DEFAULT_BITS + MS_MAX_BITS) {
    // three or more levels anyway: the narrowest level-1 digit (>= 5 bits)
    // that adds no level, since every level-1 digit keeps a partly written
    // line open per block (k = 64, 2.9e9 instances, 2^21 buckets: level-1
    // scatter 33.8 ms at 8 bits, 13.8 ms at 5)
    const int further = (btot - L1_DEFAULT_BITS + MS_MAX_BITS - 1) / MS_MAX_BITS;
    d1 = std::max(5, btot - further * MS_MAX_BITS);
  } else {
    d1 = btot < L1_DEFAULT_BITS ? btot : L1_DEFAULT_BITS;
  }
  const uint64_t B1 = 1ull << d1;
  // reads per block: up to 8 chunks of L1_NT, fewer for small inputs so that
  // the grid still spreads over the CUs
  const int rpb = L1_NT * (tun.l1_chunks > 0 && tun.l1_chunks <= 64
                               ? tun.l1_chunks
                               : (int)std::min<uint64_t>(8, std::max<uint64_t>(1, R / (L1_NT * 1024ull))));
  const uint64_t nblk = (R + rpb - 1) / rpb;
  // LDS staging: room for one record per thread of the input's mean size,
  // + 25 %, for each of the two level-1 kernels
  auto staging = [&](int nt, int cap) {
    const double want = (double)nt * ((double)n / (double)R) * 1.25;
    const int b = (int)std::min<double>(cap, std::max<double>(L1_SB_MIN, want));
    return (b + 1023) & ~1023;
  };
  const int sb = staging(L1_NT, L1_SB_MAX);
  constexpr int snt = L1Scatter<Key, ENC, OWN>::NT;
  const int sbs = staging(snt, L1Scatter<Key, ENC, OWN>::SB_MAX);
  ull *boff, *O;
  uint32_t* H;
  Key* S2;
  PK_TRY(need(c, B_BOFF, B1 + 1, &boff));
  PK_TRY(need(c, B_HIST1, B1 * nblk, &H));
  PK_TRY(need(c, B_OFF1, B1 * nblk + 1, &O));
  PK_TRY(need(c, B_S, N, &S));
  PK_TRY(need(c, B_S2, N, &S2));
  PK_TRY(need(c, B_C, N, &C));
  c->timing.bucket_bits = btot;
  // fold the second level's histogram into this one when two or more levels
  // follow level 1 and its (bucket, digit) counters fit 32 KB of LDS
  constexpr int sgf = L1Scatter<Key, ENC, OWN>::GF;
  L2Fold f2{0, sgf, 0, nullptr};
  {
    const int left = btot - d1;
    const int nlev = (left + MS_MAX_BITS - 1) / MS_MAX_BITS;
    const int db = nlev > 0 ? (left + nlev - 1) / nlev : 0;
    if (L1Scatter<Key, ENC, OWN>::main && nlev >= 2 && (B1 << db) <= L2_FOLD_MAX && tun.ms_fold != 2) {
      f2.d2 = db;
      // segments of about four level-2 tiles on average (whole scatter
      // blocks): a segment's last tile is a partial one
      constexpr uint64_t TQ2 = (uint64_t)MS_NT * TileCfg<Key>::Q;
      const uint64_t want = (nblk * B1 * 4 * TQ2 + N - 1) / N;
      const uint64_t g = std::max<uint64_t>(1, (want + sgf - 1) / sgf);
      f2.spb = (int)std::min<uint64_t>(g * sgf, (nblk + sgf - 1) / sgf * sgf);
      f2.nsb = (nblk + f2.spb - 1) / f2.spb;
      const uint64_t nh2 = (B1 << db) * f2.nsb;
      PK_TRY(need(c, B_SEGH, nh2 + 2, &f2.H2));
      PK_TRY(fill_u64(c, reinterpret_cast<ull*>(f2.H2), 0, (nh2 + 1) / 2));
    }
  }
  PK_HIP(c, hipEventRecord(c->ev[2], st));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_l1_hist<Key, ENC, OWN>), dim3((unsigned)nblk), dim3(L1_NT),
                     l1_lds_bytes(sb, (int)B1, 4) + (f2.d2 ? ((uint32_t)B1 << f2.d2) * 4u : 0u), st, fa, n, rstart,
                     rlen, R, K, KB, mask, d1, rpb, sb, cf, H, f2);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[3], st));
  PK_TRY(scan_excl<uint32_t>(c, H, O, B1 * nblk));
  if (OWN) {
    // this rank's instances
    ull no = 0;
    PK_TRY(d2h(c, &no, O + B1 * nblk, 8));
    N = no;
    c->N = N;
    c->timing.n_instances = N;
    if (N == 0) {
      c->U = 0;
      PK_TRY(need(c, B_TK, 1, &TK));
      PK_TRY(need(c, B_TC, 1, &TC));
      c->have_table = true;
      return PK_OK;
    }
  }
  hipLaunchKernelGGL(k_l1_boff, dim3(grid_for(B1 + 1, 256)), dim3(256), 0, st, O, (int)B1, nblk, (ull)N,
                     boff);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[4], st));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_l1_scatter<Key, ENC, OWN>), dim3((unsigned)((nblk + sgf - 1) / sgf)),
                     dim3(snt), l1_lds_bytes(sbs, (int)B1, 12), st, fa, n, rstart, rlen, R, K, KB, mask, d1, rpb,
                     sbs, cf, O, nblk, S);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[16], st));
  if (!f2.d2) return count_partitioned<Key>(c, S, S2, C, boff, B1, d1, btot, KB, N);
  // level 2 from the folded counts: scan them in (bucket, digit, segment)
  // order, then the segment scatter
  const int db = f2.d2;
  const uint64_t B2 = B1 << db, nh2 = B2 * f2.nsb;
  ull *O2, *O2T, *boff2;
  PK_TRY(need(c, B_SEGO, nh2 + 1, &O2));
  PK_TRY(need(c, B_SEGT, nh2, &O2T));
  PK_TRY(need(c, B_BOFF2, B2 + 1, &boff2));
  PK_HIP(c, hipEventRecord(c->ev[20], st));
  PK_TRY(scan_excl<uint32_t>(c, f2.H2, O2, nh2));
  hipLaunchKernelGGL(k_seg_transpose, dim3((unsigned)((f2.nsb + 31) / 32), (unsigned)(((1u << db) + 31) / 32), (unsigned)B1),
                     dim3(256), 0, st, O2, f2.nsb, db, O2T);
  PK_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(k_seg_boff, dim3(grid_for(B2 + 1, 256)), dim3(256), 0, st, O2T, f2.nsb, db, (ull)B2, (ull)N,
                     boff2);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[21], st));
  PK_HIP(c, hipEventRecord(c->ev[22], st));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ms_scatter_seg<Key, MS_P_NT>), dim3(2u * (unsigned)c->ncu), dim3(MS_P_NT), 0, st,
                     S, O, nblk, f2.spb, f2.nsb, (int)B1, (ull)N, KB - d1 - db, db, O2T, S2);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[23], st));
  c->timing.ms_scatter_form = 3;
  c->ms_form[0] = 3;
  c->ms_levels = 1;
  return count_partitioned<Key>(c, S2, S, C, boff2, B2, d1 + db, btot, KB, N, 1);
}

// Distinct klen-byte windows of the implicit builder counters: keys from
// k_win_keys (KB bits, even), then the partition and count from memory.
template <class Key>
int windows_count_impl(pk_ctx* c, const uint8_t* fa, const ull* rstart, const uint32_t* rlen, const ull* woff,
                       uint64_t R, uint64_t N, int klen, int bits, const uint8_t* lut, int has_nul, int KB) {
  hipStream_t st = c->stream;
  c->R = R;
  c->N = N;
  c->k = klen;
  c->words = sizeof(Key) / 8;
  c->key_bits = -1;  // byte-rank codes: no macro nodes on this table
  c->timing.n_instances = N;
  c->timing.n_reads = R;
  c->ms_levels = 0;
  PK_HIP(c, hipEventRecord(c->ev[1], st));
  if (N == 0) {
    Key* TK;
    int32_t* TC;
    c->U = 0;
    PK_TRY(need(c, B_TK, 1, &TK));
    PK_TRY(need(c, B_TC, 1, &TC));
    c->have_table = true;
    return PK_OK;
  }
  Key *S, *S2;
  int32_t* C;
  ull *boff, *small;
  PK_TRY(need(c, B_S, N, &S));
  PK_TRY(need(c, B_S2, N, &S2));
  PK_TRY(need(c, B_C, N, &C));
  PK_TRY(need(c, B_BOFF, 2, &boff));
  PK_TRY(need(c, B_SMALL, 16, &small));
  PK_TRY(fill_u64(c, small + 4, 0, 1));
  PK_TRY(fill_u64(c, boff, 0, 1));
  PK_TRY(fill_u64(c, boff + 1, N, 1));
  PK_HIP(c, hipEventRecord(c->ev[2], st));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_win_keys<Key>), dim3(grid_for(R, 256)), dim3(256), 0, st, fa, rstart, rlen,
                     woff, R, klen, bits, lut, has_nul, S);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[3], st));
  PK_HIP(c, hipEventRecord(c->ev[4], st));
  PK_HIP(c, hipEventRecord(c->ev[16], st));
  const float bf = c->tuning.bucket_factor > 0.f ? c->tuning.bucket_factor : 3.0f;
  const int btot = choose_bits(N, KB / 2, HashCfg<Key>::LIMIT, bf);
  c->timing.bucket_bits = btot;
  return count_partitioned<Key>(c, S, S2, C, boff, 1, 0, btot, KB, N);
}

template <class Key>
int count_impl(pk_ctx* c, const uint8_t* fa, uint64_t n, int K) {
  c->have_table = c->have_kept = c->have_nodes = false;
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = c->pend_cutoff = c->pend_mn = c->pend_norm = false;
  c->key_bits = 2 * K;
  PK_HIP(c, hipEventRecord(c->ev[0], c->stream));
  ull* rstart;
  uint32_t* rlen;
  uint64_t R, N;
  PK_TRY(records_main(c, fa, n, K, &rstart, &rlen, &R, &N));
  return count_records<Key, ENC_MAIN>(c, fa, n, rstart, rlen, R, N, K, 2 * K, key_mask<Key>(2 * K));
}

// L-mer frequencies of the whole input into freq (device, LMER_BINS u64),
// with the record rule of Sliding_window_l (a record needs L bytes).
int lmer_frequency_impl(pk_ctx* c, const uint8_t* fa, uint64_t n, ull** freq_out) {
  ull* rs;
  uint32_t* rl;
  uint64_t R, N;
  PK_TRY(records_main(c, fa, n, LMER_L, &rs, &rl, &R, &N));
  ull* freq;
  PK_TRY(need(c, B_LFREQ, LMER_BINS, &freq));
  PK_TRY(fill_u64(c, freq, 0, LMER_BINS));
  if (R) {
    hipLaunchKernelGGL(k_lmer_freq, dim3((unsigned)std::min<uint64_t>(grid_for(R, 256), 4096)), dim3(256), 0,
                       c->stream, fa, rs, rl, R, freq);
    PK_HIP(c, hipGetLastError());
  }
  *freq_out = freq;
  return PK_OK;
}

// One rank's share of the counted table at P ranks (the table rank `rank`
// holds after the reference's all-to-all): the k-mers whose owner is rank.
template <class Key>
int rank_share_impl(pk_ctx* c, const uint8_t* fa, uint64_t n, int K, int P, int rank) {
  c->have_table = c->have_kept = c->have_nodes = false;
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = c->pend_cutoff = c->pend_mn = c->pend_norm = false;
  c->key_bits = 2 * K;
  PK_HIP(c, hipEventRecord(c->ev[0], c->stream));
  ull* freq;
  PK_TRY(lmer_frequency_impl(c, fa, n, &freq));
  ull* rstart;
  uint32_t* rlen;
  uint64_t R, N;
  PK_TRY(records_main(c, fa, n, K, &rstart, &rlen, &R, &N));
  return count_records<Key, ENC_MAIN, true>(c, fa, n, rstart, rlen, R, N, K, 2 * K, key_mask<Key>(2 * K),
                                            OwnCfg{freq, P, rank, 0});
}

// One rank's share when every k-mer belongs to rank retrieve_proc_id(kmer)
// (the -p counters, kmer_counting.cpp:1825-2091, 2095-2271).
template <class Key>
int rank_share_key_impl(pk_ctx* c, const uint8_t* fa, uint64_t n, int K, int P, int rank) {
  c->have_table = c->have_kept = c->have_nodes = false;
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = c->pend_cutoff = c->pend_mn = c->pend_norm = false;
  c->key_bits = 2 * K;
  PK_HIP(c, hipEventRecord(c->ev[0], c->stream));
  ull* rstart;
  uint32_t* rlen;
  uint64_t R, N;
  PK_TRY(records_main(c, fa, n, K, &rstart, &rlen, &R, &N));
  return count_records<Key, ENC_MAIN, true>(c, fa, n, rstart, rlen, R, N, K, 2 * K, key_mask<Key>(2 * K),
                                            OwnCfg{nullptr, P, rank, 1});
}

template <class Key>
int builder_count_impl(pk_ctx* c, const uint8_t* fa, uint64_t n, int klen, uint64_t read_len,
                       BldSyms sy) {
  c->have_table = c->have_kept = c->have_nodes = false;
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = c->pend_cutoff = c->pend_mn = c->pend_norm = false;
  c->key_bits = (int)sizeof(Key) * 8;
  PK_HIP(c, hipEventRecord(c->ev[0], c->stream));
  ull* rstart;
  uint32_t* rlen;
  uint64_t R, N;
  PK_TRY(records_builder(c, fa, n, klen, read_len, sy, &rstart, &rlen, &R, &N));
  return count_records<Key, ENC_BUILDER>(c, fa, n, rstart, rlen, R, N, klen, (int)sizeof(Key) * 8,
                                         ~Key(0));
}

__global__ void k_set_int(int* p, int v) { *p = v; }

// fixed_mb > 0: keep count >= fixed_mb with no histogram (the PaK-graph
// builder's threshold rule); else the histogram cutoff with t buckets.
template <class Key>
int cutoff_impl(pk_ctx* c, int t, uint64_t* hist_out, int* mb_out, uint64_t* nk_out, int fixed_mb = 0) {
  hipStream_t st = c->stream;
  PK_HIP(c, hipEventRecord(c->ev[12], st));
  Key* TK = reinterpret_cast<Key*>(c->buf[B_TK].p);
  int32_t* TC = reinterpret_cast<int32_t*>(c->buf[B_TC].p);
  const uint64_t U = c->U;
  ull* hist;
  ull* small;
  PK_TRY(need(c, B_HIST, (uint64_t)t + 8, &hist));
  PK_TRY(need(c, B_SMALL, 16, &small));
  PK_TRY(fill_u64(c, hist, 0, (uint64_t)t));
  const uint64_t nb = (U + CUT_PER - 1) / CUT_PER;
  int* mbd = reinterpret_cast<int*>(small + 8);
  ull *bs, *boffs;
  PK_TRY(need(c, B_KBS, nb + 1, &bs));
  PK_TRY(need(c, B_KBOFF, nb + 1, &boffs));
  if (fixed_mb > 0) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, mbd, fixed_mb);
    PK_HIP(c, hipGetLastError());
    if (nb) {
      hipLaunchKernelGGL(k_kept_count, dim3((unsigned)nb), dim3(CUT_NT), 0, st, TC, U, mbd, bs);
      PK_HIP(c, hipGetLastError());
    }
  } else if (t <= CUT_SMALL_T) {
    uint32_t* BH;
    PK_TRY(need(c, B_CUTBH, nb * (uint64_t)t + 1, &BH));
    if (nb) {
      hipLaunchKernelGGL(k_count_hist_blocks, dim3((unsigned)nb), dim3(CUT_NT), 0, st, TC, U, t, BH);
      PK_HIP(c, hipGetLastError());
      hipLaunchKernelGGL(k_hist_columns, dim3((unsigned)t), dim3(1024), 0, st, BH, nb, t, hist);
      PK_HIP(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_min_bucket, dim3(1), dim3(1), 0, st, hist, t, mbd);
    PK_HIP(c, hipGetLastError());
    if (nb) {
      hipLaunchKernelGGL(k_kept_from_hist, dim3(grid_for(nb, 256)), dim3(256), 0, st, BH, nb, U, t,
                         mbd, bs);
      PK_HIP(c, hipGetLastError());
    }
  } else {
    if (nb) {
      hipLaunchKernelGGL(k_count_hist, dim3((unsigned)nb), dim3(CUT_NT), 0, st, TC, U, t, hist);
      PK_HIP(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_min_bucket, dim3(1), dim3(1), 0, st, hist, t, mbd);
    PK_HIP(c, hipGetLastError());
    if (nb) {
      hipLaunchKernelGGL(k_kept_count, dim3((unsigned)nb), dim3(CUT_NT), 0, st, TC, U, mbd, bs);
      PK_HIP(c, hipGetLastError());
    }
  }
  PK_TRY(scan_excl<ull>(c, bs, boffs, nb));
  ull V = 0;
  int mb = 0;
  PK_HIP(c, hipMemcpyAsync(&V, boffs + nb, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipMemcpyAsync(&mb, mbd, 4, hipMemcpyDeviceToHost, st));
  if (hist_out && fixed_mb <= 0)
    PK_HIP(c, hipMemcpyAsync(hist_out, hist, (size_t)t * 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipStreamSynchronize(st));
  Key* KK;
  int32_t* KC;
  PK_TRY(need(c, B_KK, V, &KK));
  PK_TRY(need(c, B_KC2, V, &KC));
  if (nb && V) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_kept_write<Key>), dim3((unsigned)nb), dim3(CUT_NT), 0, st,
                       TK, TC, U, mbd, boffs, KK, KC);
    PK_HIP(c, hipGetLastError());
  }
  PK_HIP(c, hipEventRecord(c->ev[13], st));
  c->pend_cutoff = true;
  c->V = V;
  c->t = t;
  c->mb = mb;
  c->have_kept = true;
  c->have_nodes = false;
  *mb_out = mb;
  *nk_out = V;
  if (V == 0 && fixed_mb <= 0)
    return c->fail(PK_EEMPTY, "no k-mer has count >= min_bucket(" + std::to_string(mb) + ")");
  return PK_OK;
}

template <class Key>
int mn_impl(pk_ctx* c) {
  hipStream_t st = c->stream;
  PK_HIP(c, hipEventRecord(c->ev[14], st));
  const Key* KK = reinterpret_cast<const Key*>(c->buf[B_KK].p);
  const int32_t* KC = reinterpret_cast<const int32_t*>(c->buf[B_KC2].p);
  const uint64_t V = c->V;
  const int K = c->k;
  const uint64_t G = (V + MN_PC - 1) / MN_PC;
  ull *q, *small, *ng, *noff;
  MnPlan* plan;
  MnRec *NT, *nodes;
  uint32_t* gflags;
  PK_TRY(need(c, B_MNQ, 8, &q));
  PK_TRY(need(c, B_SMALL, 16, &small));
  PK_TRY(need(c, B_MNPLAN, G, &plan));
  PK_TRY(need(c, B_MNNT, 2 * V, &NT));
  PK_TRY(need(c, B_MNNG, G, &ng));
  PK_TRY(need(c, B_MNNOFF, G + 1, &noff));
  PK_TRY(need(c, B_MNFLAG, 2 * V, &gflags));
  ull* maxe = small + 10;
  PK_TRY(fill_u64(c, maxe, 0, 1));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mn_quarters<Key>), dim3(1), dim3(64), 0, st, KK, V, K, q);
  PK_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mn_plan<Key>), dim3(grid_for(G, 256)), dim3(256), 0, st, KK,
                     V, K, q, G, plan, maxe);
  PK_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mn_build<Key, false>), dim3((unsigned)G), dim3(MN_NT),
                     mn_lds_bytes<Key>(), st, KK, KC, K, plan, G, NT, ng, gflags);
  PK_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mn_build<Key, true>), dim3((unsigned)G), dim3(MN_NT), 0, st,
                     KK, KC, K, plan, G, NT, ng, gflags);
  PK_HIP(c, hipGetLastError());
  PK_TRY(scan_excl<ull>(c, ng, noff, G));
  ull M = 0;
  PK_TRY(d2h(c, &M, noff + G, 8));
  PK_TRY(need(c, B_NODES, M, &nodes));
  hipLaunchKernelGGL(k_mn_compact, dim3((unsigned)G), dim3(256), 0, st, NT, plan, ng, noff, G, nodes);
  PK_HIP(c, hipGetLastError());
  PK_HIP(c, hipEventRecord(c->ev[15], st));
  c->pend_mn = true;
  c->M = M;
  c->have_nodes = true;
  return PK_OK;
}

// Keeps the rank's share of the node table (in place, key order kept).
template <class Key>
int node_share_impl(pk_ctx* c, int P, int rank) {
  hipStream_t st = c->stream;
  const uint64_t M = c->M;
  MnRec* nodes = reinterpret_cast<MnRec*>(c->buf[B_NODES].p);
  uint32_t* flag;
  ull* pos;
  MnRec* tmp;
  PK_TRY(need(c, B_TU, M + 1, &flag));
  PK_TRY(need(c, B_UOFF, M + 2, &pos));
  PK_TRY(need(c, B_MNNT, M + 1, &tmp));
  if (M) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_node_share_flags<Key>), dim3(grid_for(M, 256)), dim3(256), 0, st, nodes,
                       M, P, rank, flag);
    PK_HIP(c, hipGetLastError());
  }
  PK_TRY(scan_excl<uint32_t>(c, flag, pos, M));
  ull M2 = 0;
  PK_TRY(d2h(c, &M2, pos + M, 8));
  if (M) {
    hipLaunchKernelGGL(k_node_share_write, dim3(grid_for(M, 256)), dim3(256), 0, st, nodes, M, flag, pos, tmp);
    PK_HIP(c, hipGetLastError());
  }
  if (M2) PK_HIP(c, hipMemcpyAsync(nodes, tmp, M2 * sizeof(MnRec), hipMemcpyDeviceToDevice, st));
  c->M = M2;
  return PK_OK;
}

}  // namespace

// Host input to the device through pinned staging: PIN_THREADS host threads
// each take a contiguous slice and copy it PIN_CHUNK bytes at a time into
// one of their two pinned buffers, then queue the DMA copy on the context
// stream; a buffer is refilled once the event after its last copy has
// completed.  A DMA copy from pageable memory runs at about half the rate of
// one from pinned memory (27 against 56 GB/s measured on an MI355X host), so
// the staging pays for its host memcpy.  Small inputs and pinned inputs take
// one direct copy.
constexpr int PIN_THREADS = 4;
constexpr uint64_t PIN_CHUNK = 8ull << 20;
constexpr uint64_t PIN_MIN = 16ull << 20;

// Pinned staging shared by host-memory input and file input: fill(buf, off,
// len) puts bytes [off, off + len) of the input into the pinned buffer
// (returns false on a read error).
template <class Fill>
static int h2d_pinned(pk_ctx* c, uint8_t* dst, uint64_t n, Fill&& fill, bool* read_failed);
static bool pin_buffers(pk_ctx* c);

static int h2d_staged(pk_ctx* c, uint8_t* dst, const char* src, uint64_t n) {
  // pinned (page-locked) input is copied as it is
  hipPointerAttribute_t attr{};
  const bool pinned = hipPointerGetAttributes(&attr, src) == hipSuccess && attr.type == hipMemoryTypeHost;
  (void)hipGetLastError();  // an unregistered pointer may leave an error behind
  if (n < PIN_MIN || pinned) {
    PK_HIP(c, hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream));
    return PK_OK;
  }
  if (!pin_buffers(c)) {
    // no pinned memory to be had: one direct copy
    PK_HIP(c, hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream));
    return PK_OK;
  }
  bool bad = false;
  return h2d_pinned(
      c, dst, n,
      [src](void* buf, uint64_t off, uint64_t len) {
        std::memcpy(buf, src + off, len);
        return true;
      },
      &bad);
}

static bool pin_buffers(pk_ctx* c) {
  bool ready = true;
  for (int i = 0; i < 2 * PIN_THREADS && ready; i++) {
    if (!c->pin[i] && hipHostMalloc(&c->pin[i], PIN_CHUNK, hipHostMallocDefault) != hipSuccess) {
      c->pin[i] = nullptr;
      ready = false;
    }
    if (ready && !c->pin_ev[i] && hipEventCreateWithFlags(&c->pin_ev[i], hipEventDisableTiming) != hipSuccess) {
      c->pin_ev[i] = nullptr;
      ready = false;
    }
  }
  if (!ready) (void)hipGetLastError();
  return ready;
}

template <class Fill>
static int h2d_pinned(pk_ctx* c, uint8_t* dst, uint64_t n, Fill&& fill, bool* read_failed) {
  *read_failed = false;
  if (!pin_buffers(c)) return c->fail(PK_ENOMEM, "pinned staging buffers could not be allocated");
  const uint64_t slice = ((n + PIN_THREADS - 1) / PIN_THREADS + 63) & ~63ull;
  hipError_t err[PIN_THREADS];
  bool rerr[PIN_THREADS];
  auto work = [&](int t) {
    err[t] = hipSuccess;
    rerr[t] = false;
    const uint64_t lo = std::min<uint64_t>(n, (uint64_t)t * slice), hi = std::min<uint64_t>(n, lo + slice);
    hipError_t e = hipSetDevice(c->device);
    int k = 0;
    for (uint64_t off = lo; off < hi && e == hipSuccess; off += PIN_CHUNK, k ^= 1) {
      const int b = 2 * t + k;
      const uint64_t len = std::min<uint64_t>(PIN_CHUNK, hi - off);
      if (c->pin_used[b]) e = hipEventSynchronize(c->pin_ev[b]);
      if (e != hipSuccess) break;
      if (!fill(c->pin[b], off, len)) {
        rerr[t] = true;
        break;
      }
      e = hipMemcpyAsync(dst + off, c->pin[b], len, hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess) e = hipEventRecord(c->pin_ev[b], c->stream);
      c->pin_used[b] = true;
    }
    err[t] = e;
  };
  std::vector<std::thread> th;
  for (int t = 1; t < PIN_THREADS; t++) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  for (int t = 0; t < PIN_THREADS; t++) {
    if (rerr[t]) *read_failed = true;
    if (err[t] != hipSuccess) return c->hip_fail(err[t], "staged host-to-device copy");
  }
  return PK_OK;
}

// ---- File input (pk_count_file) ----------------------------------------------
// The reference reads its rank's byte range of the FASTA file with MPI-IO
// (DivideReads, src/input_reading.cpp:69-243); one GPU takes the whole file:
// PIN_THREADS host threads pread their slices chunk by chunk into pinned
// buffers while the DMA engine drains the previous chunk into HBM.
static int file_to_device(pk_ctx* c, const char* path, const uint8_t** fa, uint64_t* nbytes) {
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return c->fail(PK_EINVAL, std::string("cannot open ") + path + ": " + std::strerror(errno));
  struct stat st {};
  if (::fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
    ::close(fd);
    return c->fail(PK_EINVAL, std::string(path) + " is not a regular file");
  }
  const uint64_t n = (uint64_t)st.st_size;
  uint8_t* d;
  int rc = need(c, B_FASTA, n + 64, &d);
  bool bad = false;
  if (rc == PK_OK && n) {
    rc = h2d_pinned(
        c, d, n,
        [fd](void* buf, uint64_t off, uint64_t len) {
          uint64_t got = 0;
          while (got < len) {
            const ssize_t r = ::pread(fd, (char*)buf + got, len - got, (off_t)(off + got));
            if (r <= 0) {
              if (r < 0 && errno == EINTR) continue;
              return false;
            }
            got += (uint64_t)r;
          }
          return true;
        },
        &bad);
  }
  ::close(fd);
  if (rc) return rc;
  if (bad) return c->fail(PK_EINVAL, std::string("read error in ") + path);
  *fa = d;
  *nbytes = n;
  return PK_OK;
}

// ---------------------------------------------------------------------------
// C ABI.

extern "C" {

int pk_ctx_create(pk_ctx** out, int device, uint64_t hbm_budget_bytes) {
  if (!out) return PK_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PK_EHIP;
  if (device < 0 || device >= ndev) return PK_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return PK_EHIP;
  pk_ctx* c = new pk_ctx();
  c->device = device;
  c->budget = hbm_budget_bytes;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    c->ncu = ncu;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PK_EHIP;
  }
  for (auto& e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return PK_EHIP;
    }
  }
  *out = c;
  return PK_OK;
}

void pk_ctx_destroy(pk_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& b : c->buf)
    if (b.p) (void)(c->stream_alloc ? hipFreeAsync(b.p, c->stream) : hipFree(b.p));
  if (c->stream_alloc && c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (int i = 0; i < 8; i++) {
    if (c->pin[i]) (void)hipHostFree(c->pin[i]);
    if (c->pin_ev[i]) (void)hipEventDestroy(c->pin_ev[i]);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* pk_last_error(const pk_ctx* c) { return c ? c->err.c_str() : "null context"; }
void* pk_ctx_stream(pk_ctx* c) { return c ? (void*)c->stream : nullptr; }
uint64_t pk_ctx_device_bytes(const pk_ctx* c) { return c ? c->held : 0; }

int pk_ctx_set_tuning(pk_ctx* c, const pk_tuning* t) {
  if (!c) return PK_EINVAL;
  if (!t) {
    c->tuning = pk_tuning{};
    return PK_OK;
  }
  if (t->l1_chunks < 0 || t->l1_chunks > 64 || t->l1_bits < 0 || t->l1_bits > L1_MAX_BITS ||
      !(t->bucket_factor >= 0.f && t->bucket_factor <= 64.f) || t->ms_scatter < 0 || t->ms_scatter > 2 ||
      t->ms_fold < 0 || t->ms_fold > 2)
    return c->fail(PK_EINVAL, "tuning values out of range");
  c->tuning = *t;
  return PK_OK;
}

int pk_ctx_set_alloc_mode(pk_ctx* c, int mode) {
  if (!c || (mode != PK_ALLOC_DEVICE && mode != PK_ALLOC_STREAM_ORDERED)) return PK_EINVAL;
  const bool want = mode == PK_ALLOC_STREAM_ORDERED;
  if (want == c->stream_alloc) return PK_OK;
  PK_HIP(c, hipSetDevice(c->device));
  if (want) {
    int pools = 0;
    if (hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, c->device) != hipSuccess || !pools)
      return c->fail(PK_EINVAL, "the device has no stream-ordered memory pools");
  }
  // buffers are freed the way they were allocated
  PK_TRY(pk_ctx_release_buffers(c));
  c->stream_alloc = want;
  return PK_OK;
}

int pk_ctx_release_buffers(pk_ctx* c) {
  if (!c) return PK_EINVAL;
  PK_HIP(c, hipSetDevice(c->device));
  PK_HIP(c, hipStreamSynchronize(c->stream));
  for (auto& b : c->buf) {
    if (b.p) PK_HIP(c, c->stream_alloc ? hipFreeAsync(b.p, c->stream) : hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
  }
  if (c->stream_alloc) PK_HIP(c, hipStreamSynchronize(c->stream));
  c->held = 0;
  c->have_table = c->have_kept = c->have_nodes = false;
  c->have_norm = false;
  c->norm_n = 0;
  c->pend_count = c->pend_split = c->pend_cutoff = c->pend_mn = c->pend_norm = false;
  return PK_OK;
}

static int fasta_to_device(pk_ctx* c, const char* fasta, uint64_t nbytes, int fasta_on_device,
                           const uint8_t** fa);

int pk_count(pk_ctx* c, const char* fasta, uint64_t nbytes, int k, int fasta_on_device,
             pk_table* out) {
  if (!c || !out || (!fasta && nbytes)) return PK_EINVAL;
  // a failure before the count itself leaves no stale figures behind
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = false;
  if (k < 8 || k > 64) return c->fail(PK_EINVAL, "k must be in [8, 64]");
  PK_HIP(c, hipSetDevice(c->device));
  const uint8_t* fa;
  PK_TRY(fasta_to_device(c, fasta, nbytes, fasta_on_device, &fa));
  int rc = (k <= 32) ? count_impl<uint64_t>(c, fa, nbytes, k) : count_impl<u128>(c, fa, nbytes, k);
  if (rc) return rc;
  out->n = c->U;
  out->words = c->words;
  out->k = c->k;
  out->keys = reinterpret_cast<const uint64_t*>(c->buf[B_TK].p);
  out->counts = reinterpret_cast<const int32_t*>(c->buf[B_TC].p);
  return PK_OK;
}

// ---- File input (pk_count_file): file_to_device above the C ABI --------------

int pk_count_file(pk_ctx* c, const char* path, int k, pk_table* out) {
  if (!c || !out || !path) return PK_EINVAL;
  if (k < 8 || k > 64) return c->fail(PK_EINVAL, "k must be in [8, 64]");
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = false;
  PK_HIP(c, hipSetDevice(c->device));
  const uint8_t* fa;
  uint64_t n = 0;
  PK_TRY(file_to_device(c, path, &fa, &n));
  int rc = (k <= 32) ? count_impl<uint64_t>(c, fa, n, k) : count_impl<u128>(c, fa, n, k);
  if (rc) return rc;
  out->n = c->U;
  out->words = c->words;
  out->k = c->k;
  out->keys = reinterpret_cast<const uint64_t*>(c->buf[B_TK].p);
  out->counts = reinterpret_cast<const int32_t*>(c->buf[B_TC].p);
  return PK_OK;
}

// ---- Allocation-free count (pk_count_noalloc) --------------------------------
// Every intermediate buffer of the count lives in one file-scope device array
// per GPU; its address is taken once per device with hipGetSymbolAddress and
// only that address is handed to kernels and copies.
constexpr uint64_t NOALLOC_SCRATCH = PK_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MAX_DEVICES] = {};

static int scratch_address(pk_ctx* c, char** out) {
  if (c->device < 0 || c->device >= MAX_DEVICES) return c->fail(PK_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void* p = nullptr;
    // the calling thread has made c->device current
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return c->hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[c->device] = p;
  }
  *out = static_cast<char*>(g_scratch_ptr[c->device]);
  return PK_OK;
}

int pk_count_noalloc(pk_ctx* c, const char* fasta_dev, uint64_t nbytes, int k, uint64_t* keys_dev,
                     int32_t* counts_dev, uint64_t capacity, uint64_t* n_out) {
  if (!c || !n_out || (!fasta_dev && nbytes) || (capacity && (!keys_dev || !counts_dev))) return PK_EINVAL;
  if (k < 8 || k > 64) return c->fail(PK_EINVAL, "k must be in [8, 64]");
  *n_out = 0;
  PK_HIP(c, hipSetDevice(c->device));
  char* arena = nullptr;
  PK_TRY(scratch_address(c, &arena));
  // the scratch array is shared by every call on this device: a call still
  // running on this context's stream finishes before its contents change
  PK_HIP(c, hipStreamSynchronize(c->stream));
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = false;
  const int words = k <= 32 ? 1 : 2;
  // the context's own buffers step aside for the call
  DevBuf saved[B_COUNT];
  for (int i = 0; i < B_COUNT; i++) {
    saved[i] = c->buf[i];
    c->buf[i] = DevBuf{};
    c->arena_fixed[i] = false;
  }
  c->buf[B_TK] = DevBuf{keys_dev, capacity * (uint64_t)words * 8};
  c->buf[B_TC] = DevBuf{counts_dev, capacity * 4};
  c->arena_fixed[B_TK] = c->arena_fixed[B_TC] = true;
  c->arena = arena;
  c->arena_cap = NOALLOC_SCRATCH;
  c->arena_used = 0;
  c->arena_mode = true;
  c->U = 0;
  const uint8_t* fa = reinterpret_cast<const uint8_t*>(fasta_dev);
  int rc = (k <= 32) ? count_impl<uint64_t>(c, fa, nbytes, k) : count_impl<u128>(c, fa, nbytes, k);
  c->arena_mode = false;
  for (int i = 0; i < B_COUNT; i++) {
    c->buf[i] = saved[i];
    c->arena_fixed[i] = false;
  }
  // the table lives in the caller's arrays, not in the context
  c->have_table = c->have_kept = c->have_nodes = false;
  if (rc) return rc;
  *n_out = c->U;
  return PK_OK;
}


static int fasta_to_device(pk_ctx* c, const char* fasta, uint64_t nbytes, int fasta_on_device,
                           const uint8_t** fa) {
  if (fasta_on_device) {
    *fa = reinterpret_cast<const uint8_t*>(fasta);
    return PK_OK;
  }
  uint8_t* d;
  PK_TRY(need(c, B_FASTA, nbytes + 64, &d));
  if (nbytes) PK_TRY(h2d_staged(c, d, fasta, nbytes));
  *fa = d;
  return PK_OK;
}

// The rewrite of pk_normalize / pk_count_format; events 17-18 bracket it.
static int normalize_input(pk_ctx* c, const char* data, uint64_t nbytes, int format, int on_device,
                           const uint8_t** out, uint64_t* out_n) {
  if (format != PK_FMT_FASTA_STRICT && format != PK_FMT_FASTA && format != PK_FMT_FASTQ)
    return c->fail(PK_EINVAL, "unknown input format " + std::to_string(format));
  const uint8_t* in;
  PK_TRY(fasta_to_device(c, data, nbytes, on_device, &in));
  c->have_norm = false;
  PK_HIP(c, hipEventRecord(c->ev[17], c->stream));
  if (format == PK_FMT_FASTA_STRICT) {
    uint8_t* d;
    PK_TRY(need(c, B_NORM, nbytes + 64, &d));
    if (nbytes) PK_HIP(c, hipMemcpyAsync(d, in, nbytes, hipMemcpyDeviceToDevice, c->stream));
    *out = d;
    *out_n = nbytes;
  } else {
    PK_TRY(normalize_impl(c, in, nbytes, format, out, out_n));
  }
  PK_HIP(c, hipEventRecord(c->ev[18], c->stream));
  c->norm_n = *out_n;
  c->have_norm = true;
  return PK_OK;
}

int pk_normalize(pk_ctx* c, const char* data, uint64_t nbytes, int format, int data_on_device,
                 uint64_t* out_bytes) {
  if (!c || !out_bytes || (!data && nbytes)) return PK_EINVAL;
  PK_HIP(c, hipSetDevice(c->device));
  const uint8_t* norm;
  uint64_t nn;
  PK_TRY(normalize_input(c, data, nbytes, format, data_on_device, &norm, &nn));
  c->timing.normalize_ms = 0.f;
  c->pend_norm = true;
  *out_bytes = nn;
  return PK_OK;
}

int pk_download_normalized(pk_ctx* c, char* host_out) {
  if (!c || !host_out) return PK_EINVAL;
  if (!c->have_norm) return c->fail(PK_ESTATE, "no rewritten input");
  if (c->norm_n == 0) return PK_OK;
  return d2h(c, host_out, c->buf[B_NORM].p, c->norm_n);
}

int pk_count_format(pk_ctx* c, const char* data, uint64_t nbytes, int k, int format, int data_on_device,
                    pk_table* out) {
  if (!c || !out || (!data && nbytes)) return PK_EINVAL;
  // a failure before the count itself leaves no stale figures behind
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = false;
  if (k < 8 || k > 64) return c->fail(PK_EINVAL, "k must be in [8, 64]");
  PK_HIP(c, hipSetDevice(c->device));
  if (format == PK_FMT_FASTA_STRICT) return pk_count(c, data, nbytes, k, data_on_device, out);
  const uint8_t* norm;
  uint64_t nn;
  PK_TRY(normalize_input(c, data, nbytes, format, data_on_device, &norm, &nn));
  int rc = (k <= 32) ? count_impl<uint64_t>(c, norm, nn, k) : count_impl<u128>(c, norm, nn, k);
  if (rc) return rc;
  c->pend_norm = true;
  out->n = c->U;
  out->words = c->words;
  out->k = c->k;
  out->keys = reinterpret_cast<const uint64_t*>(c->buf[B_TK].p);
  out->counts = reinterpret_cast<const int32_t*>(c->buf[B_TC].p);
  return PK_OK;
}

int pk_count_rank_share(pk_ctx* c, const char* fasta, uint64_t nbytes, int k, int nranks, int rank,
                        int fasta_on_device, pk_table* out) {
  if (!c || !out || (!fasta && nbytes)) return PK_EINVAL;
  // a failure before the count itself leaves no stale figures behind
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = false;
  if (k < 8 || k > 64) return c->fail(PK_EINVAL, "k must be in [8, 64]");
  if (nranks < 1 || rank < 0 || rank >= nranks) return c->fail(PK_EINVAL, "need 0 <= rank < nranks");
  if (nranks == 1) return pk_count(c, fasta, nbytes, k, fasta_on_device, out);
  PK_HIP(c, hipSetDevice(c->device));
  const uint8_t* fa;
  PK_TRY(fasta_to_device(c, fasta, nbytes, fasta_on_device, &fa));
  int rc = (k <= 32) ? rank_share_impl<uint64_t>(c, fa, nbytes, k, nranks, rank)
                     : rank_share_impl<u128>(c, fa, nbytes, k, nranks, rank);
  if (rc) return rc;
  out->n = c->U;
  out->words = c->words;
  out->k = c->k;
  out->keys = reinterpret_cast<const uint64_t*>(c->buf[B_TK].p);
  out->counts = reinterpret_cast<const int32_t*>(c->buf[B_TC].p);
  return PK_OK;
}

int pk_count_rank_share_by_key(pk_ctx* c, const char* fasta, uint64_t nbytes, int k, int nranks, int rank,
                               int fasta_on_device, pk_table* out) {
  if (!c || !out || (!fasta && nbytes)) return PK_EINVAL;
  if (k < 8 || k > 64) return c->fail(PK_EINVAL, "k must be in [8, 64]");
  if (nranks < 1 || rank < 0 || rank >= nranks) return c->fail(PK_EINVAL, "need 0 <= rank < nranks");
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = false;
  if (nranks == 1) return pk_count(c, fasta, nbytes, k, fasta_on_device, out);
  PK_HIP(c, hipSetDevice(c->device));
  const uint8_t* fa;
  PK_TRY(fasta_to_device(c, fasta, nbytes, fasta_on_device, &fa));
  int rc = (k <= 32) ? rank_share_key_impl<uint64_t>(c, fa, nbytes, k, nranks, rank)
                     : rank_share_key_impl<u128>(c, fa, nbytes, k, nranks, rank);
  if (rc) return rc;
  out->n = c->U;
  out->words = c->words;
  out->k = c->k;
  out->keys = reinterpret_cast<const uint64_t*>(c->buf[B_TK].p);
  out->counts = reinterpret_cast<const int32_t*>(c->buf[B_TC].p);
  return PK_OK;
}

int pk_lmer_frequency(pk_ctx* c, const char* fasta, uint64_t nbytes, int fasta_on_device,
                      uint64_t* freq_host) {
  if (!c || !freq_host || (!fasta && nbytes)) return PK_EINVAL;
  PK_HIP(c, hipSetDevice(c->device));
  const uint8_t* fa;
  PK_TRY(fasta_to_device(c, fasta, nbytes, fasta_on_device, &fa));
  ull* freq;
  PK_TRY(lmer_frequency_impl(c, fa, nbytes, &freq));
  return d2h(c, freq_host, freq, LMER_BINS * 8);
}

int pk_builder_count(pk_ctx* c, const char* fasta, uint64_t nbytes, int klen, uint64_t read_len,
                     int kmer_bits, char delimiter, char skip_symbol, int fasta_on_device,
                     pk_table* out, uint64_t* n_reads) {
  if (!c || !out || (!fasta && nbytes)) return PK_EINVAL;
  if (kmer_bits != 64 && kmer_bits != 128) return c->fail(PK_EINVAL, "kmer_bits must be 64 or 128");
  if (klen < 1) return c->fail(PK_EINVAL, "klen must be >= 1");
  if (read_len > 0xFFFFFFF0ull) return c->fail(PK_EINVAL, "read_len must be < 2^32 - 16");
  if (delimiter == skip_symbol) return c->fail(PK_EINVAL, "delimiter and skip_symbol must differ");
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = false;
  const BldSyms sy{(uint8_t)delimiter, (uint8_t)skip_symbol};
  PK_HIP(c, hipSetDevice(c->device));
  const uint8_t* fa;
  PK_TRY(fasta_to_device(c, fasta, nbytes, fasta_on_device, &fa));
  int rc = kmer_bits == 64 ? builder_count_impl<uint64_t>(c, fa, nbytes, klen, read_len, sy)
                           : builder_count_impl<u128>(c, fa, nbytes, klen, read_len, sy);
  if (rc) return rc;
  out->n = c->U;
  out->words = c->words;
  out->k = c->k;
  out->keys = reinterpret_cast<const uint64_t*>(c->buf[B_TK].p);
  out->counts = reinterpret_cast<const int32_t*>(c->buf[B_TC].p);
  if (n_reads) *n_reads = c->R;
  return PK_OK;
}

int pk_builder_count_windows(pk_ctx* c, const char* fasta, uint64_t nbytes, int klen, uint64_t read_len,
                             char delimiter, char skip_symbol, int mode, int fasta_on_device, uint64_t* n_reads,
                             uint64_t* n_windows, uint64_t* n_distinct) {
  if (!c || !n_reads || !n_windows || !n_distinct || (!fasta && nbytes)) return PK_EINVAL;
  if (mode != PK_WINDOWS_V0 && mode != PK_WINDOWS_V1) return c->fail(PK_EINVAL, "unknown window mode");
  if (klen < 1 || klen > 128) return c->fail(PK_EINVAL, "klen must be in [1, 128]");
  if (read_len > 0xFFFFFFF0ull) return c->fail(PK_EINVAL, "read_len must be < 2^32 - 16");
  if (read_len < (uint64_t)klen)
    return c->fail(PK_EINVAL, "read_len < klen: the reference's window count wraps below zero");
  if (delimiter == skip_symbol) return c->fail(PK_EINVAL, "delimiter and skip_symbol must differ");
  c->timing = pk_timing{};
  c->pend_count = c->pend_split = c->pend_cutoff = c->pend_mn = c->pend_norm = false;
  c->have_table = c->have_kept = c->have_nodes = false;
  const BldSyms sy{(uint8_t)delimiter, (uint8_t)skip_symbol};
  PK_HIP(c, hipSetDevice(c->device));
  const uint8_t* fa;
  PK_TRY(fasta_to_device(c, fasta, nbytes, fasta_on_device, &fa));
  hipStream_t st = c->stream;
  PK_HIP(c, hipEventRecord(c->ev[0], st));
  ull* small;
  PK_TRY(need(c, B_SMALL, 16, &small));
  PK_TRY(fill_u64(c, small, ~0ull, 4));
  PK_TRY(fill_u64(c, small + 12, 0, 2));
  ull* ev;
  uint64_t m, p0;
  PK_TRY(event_index(c, fa, nbytes, sy.delim, sy.skip, small, &ev, &m, &p0));
  ull *rstart, *woff, *idx;
  uint32_t *rlen, *wc, *flag;
  uint32_t* bad = reinterpret_cast<uint32_t*>(small + 12);
  uint32_t* present = reinterpret_cast<uint32_t*>(small + 8);  // 8 words = small[8..11]
  PK_TRY(need(c, B_TU, m + 1, &flag));
  PK_TRY(need(c, B_UOFF, m + 2, &idx));
  if (m) {
    hipLaunchKernelGGL(k_win_flags, dim3(grid_for(m, 256)), dim3(256), 0, st, fa, nbytes, ev, m, klen, read_len, sy,
                       flag, bad);
    PK_HIP(c, hipGetLastError());
  }
  PK_TRY(scan_excl<uint32_t>(c, flag, idx, m));
  ull hv[2];
  PK_HIP(c, hipMemcpyAsync(&hv[0], idx + m, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipMemcpyAsync(&hv[1], bad, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipStreamSynchronize(st));
  const uint64_t R = hv[0];
  const bool cut = (hv[1] & 0xFFFFFFFFull) != 0, shorter = (hv[1] >> 32) != 0;
  // concurrent_kmer_counter.hpp:965-967 throws on a read cut short; _v1
  // sizes its window array with read_len - klen + 1 (:1427), which wraps
  // for a read shorter than klen
  if (mode == PK_WINDOWS_V0 && cut) return c->fail(PK_EFORMAT, "Variable size read");
  if (mode == PK_WINDOWS_V1 && shorter) return c->fail(PK_EFORMAT, "a read is shorter than klen");
  PK_TRY(need(c, B_RSTART, R + 1, &rstart));
  PK_TRY(need(c, B_RLEN, R + 1, &rlen));
  PK_TRY(need(c, B_KC, R + 1, &wc));
  PK_TRY(need(c, B_KOFS, R + 2, &woff));
  if (m) {
    hipLaunchKernelGGL(k_win_write, dim3(grid_for(m, 256)), dim3(256), 0, st, fa, nbytes, ev, m, klen, read_len, sy,
                       idx, rstart, rlen, wc);
    PK_HIP(c, hipGetLastError());
  }
  PK_TRY(scan_excl<uint32_t>(c, wc, woff, R));
  PK_TRY(fill_u64(c, small + 8, 0, 4));
  if (R) {
    hipLaunchKernelGGL(k_win_alpha, dim3((unsigned)std::min<uint64_t>((R + 3) / 4, 4096)), dim3(64, 4), 0, st, fa,
                       rstart, rlen, R, present);
    PK_HIP(c, hipGetLastError());
  }
  ull hp[5];
  PK_HIP(c, hipMemcpyAsync(hp, small + 8, 32, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipMemcpyAsync(&hp[4], woff + R, 8, hipMemcpyDeviceToHost, st));
  PK_HIP(c, hipStreamSynchronize(st));
  const uint64_t N = hp[4];
  // byte value -> rank among the values present
  const uint32_t* pw = reinterpret_cast<const uint32_t*>(hp);
  int nval = 0;
  for (int v = 0; v < 256; v++) {
    c->win_lut[v] = 0;
    if (pw[v >> 5] >> (v & 31) & 1u) c->win_lut[v] = (uint8_t)nval++;
  }
  int bits = 1;
  while ((1 << bits) < nval) bits++;
  int KB = klen * bits;
  if (KB > 128)
    return c->fail(PK_EINVAL, "the reads hold " + std::to_string(nval) + " byte values: " + std::to_string(klen) +
                                  " x " + std::to_string(bits) + " bits exceed a 128-bit window key");
  KB += KB & 1;
  const int has_nul = (pw[0] & 1u) ? 1 : 0;
  uint8_t* lut;
  PK_TRY(need(c, B_WORDER, 256, &lut));
  PK_HIP(c, hipMemcpyAsync(lut, c->win_lut, 256, hipMemcpyHostToDevice, st));
  int rc = KB <= 64 ? windows_count_impl<uint64_t>(c, fa, rstart, rlen, woff, R, N, klen, bits, lut, has_nul, KB)
                    : windows_count_impl<u128>(c, fa, rstart, rlen, woff, R, N, klen, bits, lut, has_nul, KB);
  if (rc) return rc;
  *n_reads = R;
  *n_windows = N;
  *n_distinct = c->U;
  return PK_OK;
}

int pk_cutoff(pk_ctx* c, int t, uint64_t* hist_t, int* min_bucket, uint64_t* n_kept) {
  if (!c || !min_bucket || !n_kept) return PK_EINVAL;
  if (!c->have_table) return c->fail(PK_ESTATE, "pk_cutoff before pk_count");
  if (t == 0) t = 21;
  if (t < 0 || t > T_MAX) return c->fail(PK_EINVAL, "t must be in [0, 4096]");
  PK_HIP(c, hipSetDevice(c->device));
  return c->words == 1 ? cutoff_impl<uint64_t>(c, t, hist_t, min_bucket, n_kept)
                       : cutoff_impl<u128>(c, t, hist_t, min_bucket, n_kept);
}

int pk_cutoff_threshold(pk_ctx* c, uint64_t threshold, uint64_t* n_kept) {
  if (!c || !n_kept) return PK_EINVAL;
  if (!c->have_table) return c->fail(PK_ESTATE, "pk_cutoff_threshold before a count");
  PK_HIP(c, hipSetDevice(c->device));
  if (threshold >= 0x7FFFFFFFull) {
    // counts are int32: no count exceeds such a threshold, the kept table is empty
    PK_HIP(c, hipStreamSynchronize(c->stream));
    int32_t* KC;
    uint64_t* KK;
    PK_TRY(need(c, B_KK, c->words, &KK));
    PK_TRY(need(c, B_KC2, 1, &KC));
    c->V = 0;
    c->t = 1;
    c->mb = 0;
    c->have_kept = true;
    c->have_nodes = false;
    c->pend_cutoff = false;
    c->timing.cutoff_ms = 0.f;
    *n_kept = 0;
    return PK_OK;
  }
  const int mb = (int)threshold + 1;
  int mbo = 0;
  return c->words == 1 ? cutoff_impl<uint64_t>(c, 1, nullptr, &mbo, n_kept, mb)
                       : cutoff_impl<u128>(c, 1, nullptr, &mbo, n_kept, mb);
}

int pk_kept_table(pk_ctx* c, pk_table* out) {
  if (!c || !out) return PK_EINVAL;
  if (!c->have_kept) return c->fail(PK_ESTATE, "no kept table");
  out->n = c->V;
  out->words = c->words;
  out->k = c->k;
  out->keys = reinterpret_cast<const uint64_t*>(c->buf[B_KK].p);
  out->counts = reinterpret_cast<const int32_t*>(c->buf[B_KC2].p);
  return PK_OK;
}

int pk_macro_nodes(pk_ctx* c, pk_mn_table* out) {
  if (!c || !out) return PK_EINVAL;
  if (!c->have_kept || c->V == 0) return c->fail(PK_ESTATE, "pk_macro_nodes needs a non-empty kept table");
  if (c->key_bits != 2 * c->k)
    return c->fail(PK_ESTATE, "macro nodes need k-mer keys of exactly 2k bits (builder tables with "
                              "klen != 4 * sizeof(kmer_t) keep extra bases)");
  PK_HIP(c, hipSetDevice(c->device));
  int rc = c->words == 1 ? mn_impl<uint64_t>(c) : mn_impl<u128>(c);
  if (rc) return rc;
  out->n = c->M;
  out->k = c->k;
  out->nodes = reinterpret_cast<const pk_mnode*>(c->buf[B_NODES].p);
  return PK_OK;
}

int pk_macro_nodes_rank_share(pk_ctx* c, int nranks, int rank, pk_mn_table* out) {
  if (!c || !out) return PK_EINVAL;
  if (nranks < 1 || rank < 0 || rank >= nranks) return c->fail(PK_EINVAL, "need 0 <= rank < nranks");
  PK_TRY(pk_macro_nodes(c, out));
  if (nranks == 1) return PK_OK;
  int rc = c->words == 1 ? node_share_impl<uint64_t>(c, nranks, rank) : node_share_impl<u128>(c, nranks, rank);
  if (rc) return rc;
  out->n = c->M;
  return PK_OK;
}

int pk_mn_wiring(pk_ctx* c, int coverage, const uint8_t* pre_order, pk_mnwire* out) {
  if (!c || !out || coverage < 1) return PK_EINVAL;
  if (!c->have_nodes) return c->fail(PK_ESTATE, "pk_mn_wiring before pk_macro_nodes");
  if (c->M == 0) return PK_OK;
  PK_HIP(c, hipSetDevice(c->device));
  uint8_t* ord = nullptr;
  if (pre_order) {
    PK_TRY(need(c, B_WORDER, c->M * 4, &ord));
    PK_HIP(c, hipMemcpyAsync(ord, pre_order, c->M * 4, hipMemcpyHostToDevice, c->stream));
  }
  pk_mnwire* w;
  PK_TRY(need(c, B_WIRE, c->M, &w));
  hipLaunchKernelGGL(k_mn_wire, dim3(grid_for(c->M, 256)), dim3(256), 0, c->stream,
                     reinterpret_cast<const MnRec*>(c->buf[B_NODES].p), c->M, ord, coverage, w);
  PK_HIP(c, hipGetLastError());
  return d2h(c, out, w, c->M * sizeof(pk_mnwire));
}

int pk_download_table(pk_ctx* c, uint64_t* keys, int32_t* counts) {
  if (!c) return PK_EINVAL;
  if (!c->have_table) return c->fail(PK_ESTATE, "no table");
  if (c->U == 0) return PK_OK;
  if (keys) PK_TRY(d2h(c, keys, c->buf[B_TK].p, c->U * c->words * 8));
  if (counts) PK_TRY(d2h(c, counts, c->buf[B_TC].p, c->U * 4));
  return PK_OK;
}

int pk_download_kept(pk_ctx* c, uint64_t* keys, int32_t* counts) {
  if (!c) return PK_EINVAL;
  if (!c->have_kept) return c->fail(PK_ESTATE, "no kept table");
  if (c->V == 0) return PK_OK;
  if (keys) PK_TRY(d2h(c, keys, c->buf[B_KK].p, c->V * c->words * 8));
  if (counts) PK_TRY(d2h(c, counts, c->buf[B_KC2].p, c->V * 4));
  return PK_OK;
}

int pk_download_nodes(pk_ctx* c, pk_mnode* nodes) {
  if (!c || !nodes) return PK_EINVAL;
  if (!c->have_nodes) return c->fail(PK_ESTATE, "no macro nodes");
  if (c->M == 0) return PK_OK;
  return d2h(c, nodes, c->buf[B_NODES].p, c->M * sizeof(pk_mnode));
}

int pk_last_timing(const pk_ctx* cc, pk_timing* out) {
  if (!cc || !out) return PK_EINVAL;
  pk_ctx* c = const_cast<pk_ctx*>(cc);
  auto el = [&](int a, int b) {
    float v = 0;
    (void)hipEventElapsedTime(&v, c->ev[a], c->ev[b]);
    return v;
  };
  if (c->pend_count || c->pend_cutoff || c->pend_mn || c->pend_norm) {
    PK_HIP(c, hipSetDevice(c->device));
    PK_HIP(c, hipStreamSynchronize(c->stream));
  }
  if (c->pend_count) {
    c->timing.parse_ms = el(0, 1);
    c->timing.hist_ms = el(2, 3);
    c->timing.scatter_ms = el(4, 5);
    c->timing.l1_scatter_ms = el(4, 16);
    c->timing.sort_ms = el(6, 7);
    c->timing.split_ms = c->pend_split ? el(8, 9) : 0.f;
    c->timing.compact_ms = el(10, 11);
    const int nl = c->ms_levels < MS_TIMED_LEVELS ? c->ms_levels : MS_TIMED_LEVELS;
    c->timing.ms_hist_ms = c->timing.ms_scatter_ms = 0.f;
    for (int l = 0; l < MS_TIMED_LEVELS; l++) {
      const bool on = l < nl;
      c->timing.ms_level_hist_ms[l] = on ? el(20 + 4 * l, 21 + 4 * l) : 0.f;
      c->timing.ms_level_scatter_ms[l] = on ? el(22 + 4 * l, 23 + 4 * l) : 0.f;
      c->timing.ms_level_form[l] = on ? c->ms_form[l] : 0;
      c->timing.ms_hist_ms += c->timing.ms_level_hist_ms[l];
      c->timing.ms_scatter_ms += c->timing.ms_level_scatter_ms[l];
    }
    c->timing.ms_levels = c->ms_levels;
    c->pend_count = false;
  }
  if (c->pend_cutoff) {
    c->timing.cutoff_ms = el(12, 13);
    c->pend_cutoff = false;
  }
  if (c->pend_mn) {
    c->timing.mn_ms = el(14, 15);
    c->pend_mn = false;
  }
  if (c->pend_norm) {
    c->timing.normalize_ms = el(17, 18);
    c->pend_norm = false;
  }
  *out = c->timing;
  return PK_OK;
}

}  // extern "C"
