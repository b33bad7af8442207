/* pakhip.h -- C ABI of the MI355X (gfx950) k-mer counting, histogram cutoff
 * and macro-node construction library (libpakhip.so).
 *
 * Plain C: pointers, sizes and integer codes only.  Every entry point
 * returns PK_OK (0) or a negative PK_E* code; pk_last_error() gives the
 * message of the last failure on a context.  Nothing throws across the ABI.
 *
 * Reference interfaces replaced (paths relative to the vkundeti/pakman_plus
 * source tree):
 *   pk_count            perform_kmer_counting(const char*, size_t)
 *                       src/kmer_counting.cpp:3351 -> global kmer_proc_buf
 *                       (src/distribute_kmers.cpp:87)
 *   pk_cutoff           construct_kmer_histogram(...)  src/kmer_histogram.cpp:72
 *   pk_macro_nodes      construct_macro_nodes(MN_map, kmers)  src/mn_construct.cpp:212
 *                       (pk_cutoff + pk_macro_nodes = begin_mnode_construction,
 *                       src/mn_construct.cpp:479)
 *   pk_builder_count    Concurrent_Kmer_Partial_Counter<kmer_t>::
 *                       default_count_and_process(processor)
 *                       src/concurrent_kmer_counter.hpp:2018-2082
 *
 * Threading: a context is used by one host thread at a time; contexts are
 * independent (each owns a HIP stream and its device memory).
 *
 * Ordering: the last kernels of a call may still run on the context's
 * stream when it returns (the sizes it reports are final).  Later calls on
 * the context, pk_download_* and pk_last_timing are ordered after them; a
 * caller that reads the device pointers by other means orders itself after
 * pk_ctx_stream() first.
 */
#ifndef PAKHIP_H
#define PAKHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PK_OK 0
#define PK_EINVAL (-1)  /* bad argument (k outside [8, 64], t < 0, ...) */
#define PK_ENOMEM (-2)  /* device allocation failed / over the context budget */
#define PK_EHIP (-3)    /* a HIP runtime call failed */
#define PK_EFORMAT (-4) /* input is not strict single-line FASTA where the
                           reference asserts (src/kmer_counting.cpp:2928) */
#define PK_EEMPTY (-5)  /* no k-mer survives the cutoff (the reference calls
                           MPI_Abort, src/kmer_histogram.cpp:189-200) */
#define PK_ESTATE (-6)  /* call out of order (e.g. cutoff before count) */
#define PK_EAGAIN (-7)  /* pk_server_poll: the job has not finished */

typedef struct pk_ctx pk_ctx;

/* Sorted (ascending unsigned key), unique k-mer table.  keys holds n*words
 * uint64 (words = 1 for k <= 32, 2 for 32 < k <= 64, stored lo then hi);
 * counts are int32 as in the reference's KmerPairs (distribute_kmers.h:702-714).
 * Pointers are device pointers owned by the context, valid until the next
 * call that rebuilds the table or pk_ctx_destroy. */
typedef struct pk_table {
  uint64_t n;
  int32_t words;
  int32_t k;
  const uint64_t* keys;
  const int32_t* counts;
} pk_table;

/* One macro node (a (k-1)-mer key with its 1-base extensions).  Index
 * b = base code (A=0 C=1 T=2 G=3, distribute_kmers.h:784-786).
 * pre_cnt[a] = count of k-mer a.key (prefix a), suf_cnt[b] = count of
 * k-mer key.b (suffix b); 0 = absent.  Nodes are ascending by key. */
typedef struct pk_mnode {
  uint64_t key_lo;
  uint64_t key_hi;
  int32_t pre_cnt[4];
  int32_t suf_cnt[4];
} pk_mnode;

typedef struct pk_mn_table {
  uint64_t n;
  int32_t k;
  const pk_mnode* nodes; /* device pointer owned by the context */
} pk_mn_table;

/* Per-kernel timings of the last pk_count / pk_cutoff / pk_macro_nodes
 * (milliseconds, HIP events recorded on the context stream right around
 * the kernel launches of each stage). */
typedef struct pk_timing {
  float parse_ms;    /* line index + read table */
  float hist_ms;     /* bucket histogram (parse pass 1) */
  float scatter_ms;  /* bucket scatter (parse pass 2) */
  float sort_ms;     /* per-bucket sort + run-length count */
  float split_ms;    /* oversize-bucket splitting (0 when none) */
  float compact_ms;  /* dense table assembly */
  float cutoff_ms;   /* histogram, min_bucket, kept-table compaction */
  float mn_ms;       /* macro-node construction */
  uint64_t n_instances; /* k-mer instances of the last pk_count */
  uint64_t n_reads;     /* reads with at least one byte of sequence */
  int32_t bucket_bits;  /* partition width used by the last pk_count */
  int32_t n_split;      /* buckets that needed splitting */
  float l1_scatter_ms;  /* first scatter level alone (one kernel, part of scatter_ms) */
  float normalize_ms;   /* input-format normalization (pk_count_format; 0 otherwise) */
  float ms_hist_ms;     /* further partition levels: their histogram kernels, summed */
  float ms_scatter_ms;  /* further partition levels: their scatter kernels, summed */
  int32_t ms_levels;    /* further partition levels of the last count (timed: first 4) */
  int32_t ms_scatter_form; /* their scatter kernel: 1 persistent blocks, 2 one block
                              per tile, 3 segments of a histogram folded into
                              level 1's (the last level's; 0 when none ran) */
  float ms_level_hist_ms[4];    /* per further level (the first 4): its histogram
                                   kernel, or with form 3 the scan of the folded counts */
  float ms_level_scatter_ms[4]; /* per further level: its scatter kernel */
  int32_t ms_level_form[4];     /* per further level: its scatter form as above */
} pk_timing;

/* Create a context on HIP device `device`.  hbm_budget_bytes = 0 means no
 * limit beyond the device's memory. */
int pk_ctx_create(pk_ctx** out, int device, uint64_t hbm_budget_bytes);
void pk_ctx_destroy(pk_ctx* ctx);
const char* pk_last_error(const pk_ctx* ctx);
/* The context's hipStream_t (as void*). */
void* pk_ctx_stream(pk_ctx* ctx);
/* Device bytes currently held by the context. */
uint64_t pk_ctx_device_bytes(const pk_ctx* ctx);
/* Launch-shape choices of the counting kernels (results do not depend on
 * them).  Zero fields keep the built-in choice.
 *   l1_chunks      reads per level-1 partition block, in chunks of 512 (1..64;
 *                  built-in: 1..8 by input size)
 *   l1_bits        level-1 radix digit width (1..PK_TUNING_L1_BITS_MAX; built-in 8,
 *                  or, when the bucket count needs three or more partition
 *                  levels, the narrowest width >= 5 that adds no level)
 *   bucket_factor  instances per final bucket as a multiple of the LDS hash
 *                  table's key capacity, i.e. the inverse hash load target
 *                  (built-in 3.0)
 *   ms_scatter     form of the further partition levels' scatter: 1 =
 *                  persistent blocks (two per CU, consecutive tiles each),
 *                  2 = one block per tile (built-in: persistent once every
 *                  block gets >= 256 tiles)
 *   ms_fold        2 = never count the second partition level inside the
 *                  level-1 histogram (built-in: counted there when two or
 *                  more levels follow level 1 and the first two digits have
 *                  at most 8192 values; 1 = the same) */
#define PK_TUNING_L1_BITS_MAX 10
typedef struct pk_tuning {
  int32_t l1_chunks;
  int32_t l1_bits;
  float bucket_factor;
  int32_t ms_scatter;
  int32_t ms_fold;
} pk_tuning;
int pk_ctx_set_tuning(pk_ctx* ctx, const pk_tuning* t);

/* How the context allocates its device buffers.  PK_ALLOC_DEVICE (the
 * default): hipMalloc / hipFree, where a buffer that grows waits for the
 * whole device.  PK_ALLOC_STREAM_ORDERED: hipMallocAsync / hipFreeAsync on the
 * context's stream from the device's default pool, so a context that grows
 * a buffer does not wait on other contexts' work (the server's slots use
 * it).  Changing the mode frees the context's buffers. */
#define PK_ALLOC_DEVICE 0
#define PK_ALLOC_STREAM_ORDERED 1
int pk_ctx_set_alloc_mode(pk_ctx* ctx, int mode);

/* Free every device buffer of the context (they grow back on the next call);
 * invalidates the table, kept table and node pointers. */
int pk_ctx_release_buffers(pk_ctx* ctx);

/* Count the k-mers of a strict single-line FASTA buffer (main path of the
 * reference, kmer_counting.cpp:2890-3261).  fasta may be a host pointer
 * (copied in) or, with fasta_on_device != 0, a device pointer that stays
 * resident.  8 <= k <= 64.  On success *out describes the table. */
int pk_count(pk_ctx* ctx, const char* fasta, uint64_t nbytes, int k, int fasta_on_device,
             pk_table* out);

/* pk_count of a FASTA file (the reference's input path: its rank's byte
 * range read with MPI-IO, DivideReads, src/input_reading.cpp:69-243; one GPU
 * takes the whole file).  Host threads read slices of the file into pinned
 * buffers while the previous chunks are copied to HBM.  A file that cannot be
 * opened or read fails with PK_EINVAL. */
int pk_count_file(pk_ctx* ctx, const char* path, int k, pk_table* out);

/* Allocation-free count.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  pk_count's table of a
 * device-resident FASTA buffer, written into the caller's device arrays
 * (keys_dev: capacity * words uint64, counts_dev: capacity int32; words = 1
 * for k <= 32, else 2); *n_out = distinct k-mers.  Every intermediate buffer
 * sits in a static scratch array of PK_NOALLOC_SCRATCH_BYTES per GPU (held
 * until the process ends), so no hipMalloc / hipFree is called.  Fails with
 * PK_ENOMEM when the input's buffers exceed the scratch (about 4 MiB of
 * 100 bp-read FASTA at k = 32) or the table exceeds capacity.  All calls on
 * one GPU share its scratch array: calls that overlap on one GPU (from
 * different contexts or threads) race; calls on one context are ordered.
 * The table's last kernels may still run on the context's stream at return.
 * The context's own table, kept table and nodes are invalidated. */
#define PK_NOALLOC_SCRATCH_BYTES (64ull << 20)
int pk_count_noalloc(pk_ctx* ctx, const char* fasta_dev, uint64_t nbytes, int k, uint64_t* keys_dev,
                     int32_t* counts_dev, uint64_t capacity, uint64_t* n_out);

/* Input formats beyond the reference's strict single-line FASTA (no
 * reference counterpart: the reference reads only that format,
 * src/input_reading.cpp:69-312 and src/kmer_counting.cpp:2920-2945).  The
 * device rewrites the input into strict single-line FASTA, then counts it as
 * pk_count does; the table is pk_count's table of the rewritten input.
 *   PK_FMT_FASTA_STRICT  the reference's format, passed through unchanged
 *   PK_FMT_FASTA         multi-line FASTA: a record is a line starting with
 *                        '>' and the lines up to the next such line, whose
 *                        concatenation is the read; lines before the first
 *                        record are ignored
 *   PK_FMT_FASTQ         four-line FASTQ records ('@' header, sequence, '+'
 *                        line, qualities); the header becomes '>' + the rest
 *                        of the line and the sequence the read
 * In both rewritten formats a '\r' just before a '\n' is dropped, and every
 * record ends with '\n'.  A FASTQ input whose line count is not a multiple of
 * 4, or whose header / separator line does not start with '@' / '+', fails
 * with PK_EFORMAT. */
#define PK_FMT_FASTA_STRICT 0
#define PK_FMT_FASTA 1
#define PK_FMT_FASTQ 2
int pk_count_format(pk_ctx* ctx, const char* data, uint64_t nbytes, int k, int format, int data_on_device,
                    pk_table* out);
/* The rewrite alone: the strict single-line FASTA stays in the context
 * (*out_bytes long) for pk_download_normalized. */
int pk_normalize(pk_ctx* ctx, const char* data, uint64_t nbytes, int format, int data_on_device,
                 uint64_t* out_bytes);
int pk_download_normalized(pk_ctx* ctx, char* host_out);

/* One rank's share of the counted table at nranks ranks: the k-mers whose
 * owner is `rank`, i.e. the kmer_proc_buf that rank holds after the
 * reference's exchange (perform_kmer_counting at P = nranks,
 * src/kmer_counting.cpp:2950-3050,3351-3458).  Owner = hash31(m) % nranks
 * (src/distribute_kmers.cpp:95-105) of the k-mer's leftmost minimum-frequency
 * 8-mer m, frequencies taken over the whole input (Sliding_window_l +
 * MPI_Allreduce, src/kmer_counting.cpp:391-438,3366).  nranks == 1 is
 * pk_count. */
int pk_count_rank_share(pk_ctx* ctx, const char* fasta, uint64_t nbytes, int k, int nranks, int rank,
                        int fasta_on_device, pk_table* out);

/* One rank's share at nranks ranks when a k-mer belongs to rank
 * retrieve_proc_id(kmer) (src/distribute_kmers.cpp:95-120): the owner rule
 * of the -p counters (src/kmer_counting.cpp:1949-1950, 2136, 2320), hash31
 * (includes/distribute_kmers.h:856-873) of the k-mer for k <= 32 and uhash31
 * (:836-854, 128-bit) for k > 32, modulo nranks.  The shares partition
 * pk_count's table.  nranks == 1 is pk_count. */
int pk_count_rank_share_by_key(pk_ctx* ctx, const char* fasta, uint64_t nbytes, int k, int nranks, int rank,
                               int fasta_on_device, pk_table* out);

/* The 8-mer frequency table of the input (4^8 uint64 entries, host array),
 * as Sliding_window_l computes it (src/kmer_counting.cpp:391-438). */
int pk_lmer_frequency(pk_ctx* ctx, const char* fasta, uint64_t nbytes, int fasta_on_device,
                      uint64_t* freq_host);

/* Builder-API counting (Concurrent_Kmer_Partial_Counter<kmer_t>::
 * default_count_and_process, src/concurrent_kmer_counter.hpp:2018-2082, with
 * the ctor arguments klen, read_len, delimiter ('\n') and skip_symbol ('>')
 * of :49-56): from each skip symbol that is the first after a delimiter (or
 * in the buffer), skip to the next delimiter and take read_len bytes (fewer
 * at the end of the buffer) as the read; reads shorter than klen are skipped.
 * delimiter != skip_symbol.
 * Bases A C T G -> 0 1 2 3, any other byte -> 0 (symbol_to_kmer, :2097-2110);
 * the key is the whole kmer_t word (kmer_bits = 64 for uint64_t/size_t, 128
 * for unsigned __int128) after shifting in every byte of the window, so it
 * keeps the last kmer_bits/2 bases.  The table holds every distinct key with
 * its count, ascending (the reference hands the same pairs to its processor
 * in hash order); *n_reads = the reference's return value (reads counted).
 * out->k = klen.  pk_cutoff applies to the table; pk_macro_nodes only when
 * klen == kmer_bits / 2. */
int pk_builder_count(pk_ctx* ctx, const char* fasta, uint64_t nbytes, int klen, uint64_t read_len,
                     int kmer_bits, char delimiter, char skip_symbol, int fasta_on_device,
                     pk_table* out, uint64_t* n_reads);

/* Window counting of the builder API's implicit counters
 * (Concurrent_Kmer_Partial_Counter::concurrent_count_and_process_omp_implicit_v0
 * and _v1, src/concurrent_kmer_counter.hpp:910-1022 and 1363-1578): the same
 * records as pk_builder_count (every record counted, whatever its length),
 * every klen-byte window of each read, windows compared as the reference's
 * strncmp(a, b, klen) does (raw bytes up to the first NUL; no 2-bit
 * packing).  *n_reads = records (the reference's return value), *n_windows =
 * windows (its permutation size), *n_distinct = distinct windows.
 * PK_WINDOWS_V0 fails with PK_EFORMAT ("Variable size read") when a read is
 * cut short by the end of the buffer (the reference throws, :965-967);
 * PK_WINDOWS_V1 takes such reads but fails with PK_EFORMAT when one is
 * shorter than klen (the reference's window count wraps there, :1427).
 * read_len < klen fails with PK_EINVAL.  Windows are keyed by the rank of
 * each byte value among those present in the reads, so klen x bits per value
 * must fit 128 bits (e.g. klen = 32 with up to 16 distinct byte values);
 * otherwise PK_EINVAL.  The context's table then holds those keys (no macro
 * nodes on it). */
#define PK_WINDOWS_V0 0
#define PK_WINDOWS_V1 1
int pk_builder_count_windows(pk_ctx* ctx, const char* fasta, uint64_t nbytes, int klen, uint64_t read_len,
                             char delimiter, char skip_symbol, int mode, int fasta_on_device, uint64_t* n_reads,
                             uint64_t* n_windows, uint64_t* n_distinct);

/* Histogram cutoff on the current table (kmer_histogram.cpp:84-140):
 * hist_t[c] (host array of t entries) = number of k-mers with count c < t;
 * *min_bucket = first argmin of hist_t[1..t-1] (t == 1 -> 1); the kept
 * table = entries with count >= *min_bucket, in table order.  t == 0 uses
 * the reference's default 21 (distribute_kmers.cpp:291-294).  Returns
 * PK_EEMPTY when nothing is kept. */
int pk_cutoff(pk_ctx* ctx, int t, uint64_t* hist_t, int* min_bucket, uint64_t* n_kept);
/* Threshold cutoff of the PaK-graph builder (Pakman_Subgraph::build_graph,
 * src/pakman_graph.hpp:592-610: k-mers with count <= threshold are skipped;
 * default_pakgraph_builder.cpp:39-48 passes t - 1): the kept table = entries
 * with count > threshold, in table order.  No histogram; an empty kept table
 * is not an error.  pk_macro_nodes then builds that builder's nodes (the
 * same (k-1)-mer records; visit counts come from visit_value). */
int pk_cutoff_threshold(pk_ctx* ctx, uint64_t threshold, uint64_t* n_kept);
/* The kept table of the last pk_cutoff / pk_cutoff_threshold. */
int pk_kept_table(pk_ctx* ctx, pk_table* out);

/* Macro nodes of the kept table (mn_construct.cpp:212-477, one rank). */
int pk_macro_nodes(pk_ctx* ctx, pk_mn_table* out);

/* Macro nodes of the kept table as rank `rank` holds them at nranks ranks
 * (construct_macro_nodes, src/mn_construct.cpp:212-477): the nodes whose
 * (k-1)-mer key the rank owns, retrieve_proc_id(key) % nranks (hash31 for
 * k <= 32, uhash31 for k > 32; mn_construct.cpp:234-247), each with the
 * records pk_macro_nodes gives it (the regroup sends a node's owner every
 * kept k-mer of the node, src/kmer_histogram.cpp:140-170).  The context's
 * node table becomes the share. */
int pk_macro_nodes_rank_share(pk_ctx* ctx, int nranks, int rank, pk_mn_table* out);

/* Initial wiring of the macro nodes (initiate_mnode_wiring, src/wire.cpp:
 * 66-121, and MacroNode::setup_wiring, includes/distribute_kmers.h:550-656):
 * per node the prefix list (pre_order, or ascending base when pre_order is
 * NULL) and the suffix list (ascending base) of single-base entries with
 * (count, visit = ceil(count / coverage)), each followed by a terminal; the
 * terminal of the lighter side gets the difference of the visit totals; then
 * prefixes and suffixes in descending (visit, count) order are wired.
 * pre_order: host array of 4 bases per node (the prefixes' order, 255 =
 * unused), e.g. from pakhip::reference_prefix_order to match a reference
 * run.  out: host array of one record per node of the last pk_macro_nodes. */
typedef struct pk_wire_info {
  int32_t suffix_id;
  int32_t offset_in_suffix;
  int32_t count;
} pk_wire_info;

typedef struct pk_mnwire {
  int32_t n_prefix;  /* including the terminal */
  int32_t n_suffix;  /* including the terminal */
  int32_t n_wire;    /* filled entries of wire */
  int32_t pre_terminal_count[2];  /* (count, visit) of the prefix terminal */
  int32_t suf_terminal_count[2];  /* (count, visit) of the suffix terminal */
  pk_wire_info wire[11];          /* n_prefix + n_suffix + 1 entries in the reference */
  int32_t prefix_begin[5][2];     /* per prefix: (first wire, number of wires) */
} pk_mnwire;

int pk_mn_wiring(pk_ctx* ctx, int coverage, const uint8_t* pre_order, pk_mnwire* out);

/* Copy the current table / kept table / macro nodes to host memory.
 * Arrays must hold n*words keys, n counts, n nodes. */
int pk_download_table(pk_ctx* ctx, uint64_t* keys, int32_t* counts);
int pk_download_kept(pk_ctx* ctx, uint64_t* keys, int32_t* counts);
int pk_download_nodes(pk_ctx* ctx, pk_mnode* nodes);

int pk_last_timing(const pk_ctx* ctx, pk_timing* out);

/* ---- Batched and served forms (no single counterpart in the reference: many
 * independent read sets, each with the pk_count / pk_cutoff / pk_macro_nodes
 * semantics, overlapped on one device) ----
 *
 * A server owns n_slots contexts on one device, each driven by its own host
 * thread and HIP stream, so that the jobs' kernels overlap on the GPU.  Jobs
 * wait in a FIFO queue and start when a slot is free and their estimated HBM
 * footprint, added to the running jobs' estimates and the resident results,
 * fits the budget (a job always starts when nothing runs).  A job's results
 * (table, kept table, macro nodes) stay resident in device memory owned by
 * the server until pk_server_release.  Results are identical to the same job
 * run alone on a context.  All pk_server_* calls are thread-safe. */
typedef struct pk_job {
  const char* fasta;       /* host or device bytes, valid until the job ends */
  uint64_t nbytes;
  int32_t k;               /* 8 <= k <= 64 */
  int32_t t;               /* cutoff buckets as pk_cutoff (0 -> 21); < 0: count only */
  int32_t fasta_on_device; /* fasta is a device pointer */
  int32_t want_nodes;      /* also build macro nodes (needs t >= 0) */
} pk_job;

typedef struct pk_job_result {
  int32_t status;          /* PK_OK, or the job's PK_E* code (PK_EEMPTY keeps the table) */
  int32_t min_bucket;
  uint64_t n_reads;
  uint64_t n_instances;
  pk_table table;          /* device pointers owned by the server */
  pk_table kept;
  pk_mn_table nodes;
  double t_submit, t_start, t_end; /* seconds on the pk_server_now() clock */
  uint64_t est_bytes;      /* admission estimate of the job's working set */
  char message[160];       /* error message when status != PK_OK */
} pk_job_result;

typedef struct pk_server pk_server;

/* hbm_budget_bytes = 0: 90 % of the device memory free at creation. */
int pk_server_create(pk_server** out, int device, int n_slots, uint64_t hbm_budget_bytes);
/* Drops queued jobs, waits for running ones, frees everything. */
void pk_server_destroy(pk_server* srv);
double pk_server_now(const pk_server* srv);
int pk_server_submit(pk_server* srv, const pk_job* job, uint64_t* ticket);
/* Enqueue n jobs at once, largest first (longest-processing-time order);
 * tickets[i] belongs to jobs[i]. */
int pk_server_submit_batch(pk_server* srv, const pk_job* jobs, int n, uint64_t* tickets);
/* Block until the job has finished; fill *out. */
int pk_server_wait(pk_server* srv, uint64_t ticket, pk_job_result* out);
/* PK_EAGAIN while the job is queued or running, else as pk_server_wait. */
int pk_server_poll(pk_server* srv, uint64_t ticket, pk_job_result* out);
/* Copy a finished job's results to host arrays (any pointer may be NULL):
 * table keys/counts, kept keys/counts, hist_t (t entries), nodes. */
int pk_server_download(pk_server* srv, uint64_t ticket, uint64_t* keys, int32_t* counts,
                       uint64_t* kept_keys, int32_t* kept_counts, uint64_t* hist_t,
                       pk_mnode* nodes);
/* Per-size tuning for the jobs of a server: a job of nbytes with key words
 * w (1 for k <= 32, else 2) uses the first entry (ascending max_bytes) with
 * nbytes <= max_bytes and key_words == w (or 0 = any), else the built-in
 * choices.  Copies the table; n == 0 clears it.  Entries with key_words
 * outside 0..2 or tuning values outside pk_ctx_set_tuning's ranges fail
 * with PK_EINVAL and leave the table as it was. */
typedef struct pk_tune_entry {
  uint64_t max_bytes;
  int32_t key_words;
  pk_tuning tuning;
} pk_tune_entry;
int pk_server_set_tuning(pk_server* srv, const pk_tune_entry* entries, int n);

/* Free a finished job's device results and forget the ticket. */
int pk_server_release(pk_server* srv, uint64_t ticket);
const char* pk_server_last_error(const pk_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* PAKHIP_H */
