/*
 * lsb.h -- C ABI of the MI355X-native stable LSD radix sort of 16-byte
 * {uint64 key, uint64 val} elements.
 *
 * The reference (mppf/distributed-lsb) exposes no library API: its per-rank
 * sort is the internal function
 *     void mySort(DistributedArray<SortElement>& A,
 *                 DistributedArray<SortElement>& B)      mpi/mpi_lsbsort.cpp:591-597
 *                                                        shmem/shmem_lsbsort.cpp:469-482
 * called once by main (mpi/mpi_lsbsort.cpp:734) between two barriers.  On
 * entry A holds the elements and B is same-sized scratch; on exit A holds the
 * stable sort by key and B is clobbered.  lsb_sort() below has exactly that
 * contract for one rank's block resident in HBM.
 *
 * All pointers to elements are DEVICE pointers (HBM), 16-byte aligned.
 * Everything is plain C: no C++ types, no exceptions cross this boundary.
 * Calls are stream-ordered and asynchronous unless LSB_SYNC is set.
 */
#ifndef LSB_H
#define LSB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* struct SortElement {uint64_t key; uint64_t val;}  mpi/mpi_lsbsort.cpp:30-33,
 * shmem/shmem_lsbsort.cpp:25-28.  16 bytes, key at offset 0, little-endian.
 * Ordering is by key only (shmem/shmem_lsbsort.cpp:49-51), stable. */
typedef struct lsb_elem {
  uint64_t key;
  uint64_t val;
} lsb_elem;

/* Return codes.  The reference returns no errors (MPI errors abort,
 * mpi/mpi_lsbsort.cpp:18-20; invariants are asserts); these replace its
 * asserts and its one pre-flight guard (n/R <= INT_MAX, 647-655). */
#define LSB_OK 0
#define LSB_EINVAL (-1)  /* null / misaligned / overlapping / host pointer, bad key_bits, digit_bits or tile */
#define LSB_ENOMEM (-2)  /* device allocation failed, or caller workspace too small */
#define LSB_EHIP (-3)    /* a HIP call failed; see lsb_last_hip_error() */
#define LSB_ERANGE (-4)  /* n above LSB_MAX_N */
#define LSB_EBUSY (-5)   /* serving: job queue or memory budget full */
#define LSB_ENOTFOUND (-6) /* serving: unknown job id */

/* Largest n one lsb_sort call accepts.  Per-bucket counts and the
 * look-back status words are 32-bit (31-bit count + 1 flag bit). */
#define LSB_MAX_N ((size_t)0x7FFFFFFEu)

/* lsb_config.flags */
#define LSB_SYNC 1u     /* return only after the sort finished on the stream */
#define LSB_NO_SKIP 2u  /* run every digit pass, even single-bucket ones */
#define LSB_PROFILE 4u  /* record HIP events around each kernel of this call (lsb_profile_read) */
#define LSB_GRAPH 8u    /* run the launch sequence as a captured hipGraph, cached per buffer set */
/* Key order (at most one; default: unsigned integers of key_bits bits).
 * SIGNED: two's complement of key_bits bits.  FLOAT: IEEE-754 binary64
 * (key_bits 64) or binary32 in the low 32 bits (key_bits 32), ordered by
 * value with -0 before +0 and NaNs at the ends by their bits (negative NaNs
 * first).  Elements are moved unchanged; only the order changes.  Replaces
 * the `negs` argument the reference passes to getDigit and never uses
 * (chpl/arkouda-radix-sort.chpl:37-39). */
#define LSB_KEYS_SIGNED 16u
#define LSB_KEYS_FLOAT 32u
/* Form of the sort (same results either way).  LSB_FORM_LSD: the
 * reference's order, one stable pass per digit from the least significant
 * (mpi/mpi_lsbsort.cpp:591-597).  LSB_FORM_MSD: most significant digit
 * first -- every range is partitioned on its highest varying digit until
 * its buckets fit one workgroup, which sorts them in LDS by the digits left
 * (see lsb_sort_segmented).  LSB_FORM_LSD takes precedence; an explicit
 * digit width other than 8, tile or variant, or LSB_NO_SKIP, selects LSD.
 * Neither flag: lsb_sort runs MSD for keys wider than 32 bits and
 * n >= LSB_MSD_MIN_N with automatic digit width, tile and variant, and LSD
 * otherwise; batched sorts (lsb_sort_segmented, lsb_sort_batch) run MSD for
 * their segments above one workgroup's capacity when keys are wider than 32
 * bits, and LSD otherwise (on MI355X the two forms run level for uniform
 * 32-bit keys and LSD is faster for skewed ones).  lsb_sort_form tells
 * which form lsb_sort takes. */
#define LSB_FORM_LSD 64u
#define LSB_FORM_MSD 128u
#define LSB_SEG_LSD LSB_FORM_LSD /* names of the flags in the batched sorts */
#define LSB_SEG_MSD LSB_FORM_MSD
#define LSB_MSD_MIN_N ((size_t)1 << 24)

typedef struct lsb_config {
  uint32_t key_bits;          /* sort by the low key_bits bits of key (1..64); 0 = 64.
                                 Chapel's nBits, chpl/arkouda-radix-sort.chpl:69 */
  uint32_t digit_bits;        /* digit width d: 0 = auto, or 8 or 11 (mpi_lsbsort.cpp:22
                                 uses 16; 2^16 counters per wave do not fit 160 KiB LDS) */
  uint32_t tile;              /* elements per scatter tile: 0 = auto, else a compiled
                                 tile for the digit width (see lsb_variants) */
  uint32_t variant;           /* 0 = pick by digit_bits/tile; k > 0 = compiled variant k-1
                                 of lsb_variants() exactly */
  void* stream;               /* hipStream_t to run on; NULL = default stream */
  void* workspace;            /* device scratch of >= lsb_workspace_size() bytes;
                                 NULL = library-managed, cached per (device, stream) */
  size_t workspace_bytes;
  uint32_t flags;             /* LSB_SYNC | LSB_NO_SKIP | LSB_PROFILE | LSB_GRAPH |
                                 LSB_KEYS_SIGNED | LSB_KEYS_FLOAT | LSB_FORM_* */
  uint32_t reserved1;
} lsb_config;

/* Fills *cfg with the defaults (64-bit keys, auto digit/tile, default
 * stream, library workspace, asynchronous). */
void lsb_config_init(lsb_config* cfg);

/* Stable sort of A[0..n) by key, B[0..n) is scratch; result in A.
 * Replaces mySort (mpi/mpi_lsbsort.cpp:591-597; shmem/shmem_lsbsort.cpp:
 * 469-482), an LSD sort; the form (LSB_FORM_LSD / LSB_FORM_MSD) changes the
 * order of the work, never the result.  n = 0 and n = 1 succeed without
 * touching memory.  cfg may be NULL.  Returns LSB_OK or a negative LSB_E*
 * code. */
int lsb_sort(lsb_elem* A, lsb_elem* B, size_t n, const lsb_config* cfg);

/* The form lsb_sort(n, cfg) takes: LSB_FORM_LSD or LSB_FORM_MSD, or a
 * negative LSB_E* code for a configuration lsb_sort rejects. */
int lsb_sort_form(size_t n, const lsb_config* cfg);

/* lsb_sort with no device allocation at call time, for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference, whose mySort
 * allocates its temporaries per pass (mpi/mpi_lsbsort.cpp:523-528, 550-555).
 * Same contract and results as lsb_sort, always in the LSD form (the MSD
 * form stages its range table in pinned host memory; LSB_FORM_MSD is
 * LSB_EINVAL here), with the LSD form's launch sequence; the workspace
 * is a fixed device array of lsb_noalloc_scratch_bytes() bytes that every
 * GPU the library runs on holds until the process ends; the first call on a
 * device looks up that device's copy.  Accepts n <= LSB_NOALLOC_MAX_N
 * (LSB_ERANGE above); cfg->workspace must be NULL and LSB_GRAPH /
 * LSB_PROFILE unset (LSB_EINVAL).  All calls on one device share the array:
 * calls that overlap on one device (two streams, two threads) race. */
#define LSB_NOALLOC_MAX_N ((size_t)1 << 27)
int lsb_sort_noalloc(lsb_elem* A, lsb_elem* B, size_t n, const lsb_config* cfg);
size_t lsb_noalloc_scratch_bytes(void);

/* Bytes of device workspace lsb_sort(n, cfg) needs (0 for n <= 1). */
size_t lsb_workspace_size(size_t n, const lsb_config* cfg);

/* Digit width, tile and passes of the LSD form for (n, cfg) (what
 * lsb_sort uses when lsb_sort_form says LSB_FORM_LSD); returns LSB_OK or
 * LSB_EINVAL. */
int lsb_plan(size_t n, const lsb_config* cfg, uint32_t* digit_bits, uint32_t* tile,
             uint32_t* passes);

/* Compiled scatter variants: digit width, tile, and whether the pass runs
 * as a persistent grid that prefetches its next tile.  Writes up to cap
 * entries (any pointer may be NULL) and returns how many exist. */
int lsb_variants(uint32_t* digit_bits, uint32_t* tiles, uint32_t* persistent, int cap);

/* ------------------------------------------------------------ autotuner */

/* One tuned size class: the fastest compiled scatter variant for arrays of
 * about 2^log2n elements whose keys are at most key_bits wide. */
typedef struct lsb_tune_entry {
  uint32_t log2n;
  uint32_t key_bits;
  uint32_t digit_bits;  /* of the chosen variant */
  uint32_t tile;
  uint32_t persistent;
  uint32_t variant;     /* 1-based id into lsb_variants() */
  float ms;             /* its median sort time */
  float runner_up_ms;   /* the next-best variant's time */
} lsb_tune_entry;

/* Times every compiled variant with HIP events on uniform keys of width
 * key_bits at each size class 2^lo_log2 .. 2^hi_log2, keeps the fastest per
 * class, and makes the table the automatic choice of lsb_sort and
 * lsb_sort_segmented (digit_bits = tile = variant = 0).  Needs 48 bytes x
 * 2^hi_log2 of free device memory while it runs.  Synchronous. */
int lsb_autotune(uint32_t lo_log2, uint32_t hi_log2, uint32_t key_bits, void* stream);
/* Copies up to cap entries of the current table; returns the entry count.
 * While it is empty (never tuned, or after lsb_tune_clear) the automatic
 * choice follows a built-in table: this autotuner's output on one MI355X. */
int lsb_tune_table(lsb_tune_entry* out, int cap);
/* Saves / loads / clears the table (text file, one size class per line). */
int lsb_tune_save(const char* path);
int lsb_tune_load(const char* path);
int lsb_tune_clear(void);

const char* lsb_strerror(int code);
const char* lsb_version(void);
/* hipError_t of the last failed HIP call on this thread (0 if none). */
int lsb_last_hip_error(void);

/* ------------------------------------------------------ batched segments */

/* Sorts each segment [seg_offsets[s], seg_offsets[s+1]) of A, s < nseg,
 * stably by key and independently of the others; B (same extent) is
 * scratch; results in A.  seg_offsets is a HOST array of nseg+1
 * non-decreasing element offsets; each segment holds at most LSB_MAX_N
 * elements.  One launch sequence covers the whole batch: segments of up to
 * one workgroup's capacity (8192 elements) are sorted entirely in LDS.
 * Larger ones are partitioned most significant digit first (LSB_SEG_MSD;
 * the default for keys wider than 32 bits) with persistent onesweep passes
 * whose look-back stays inside each range, until every bucket fits one
 * workgroup and is finished in LDS; with LSB_SEG_LSD (the default for keys
 * of at most 32 bits), an explicit tile, a digit width other than 8 or
 * LSB_NO_SKIP they run LSD passes over every digit instead.  Both give each segment its stable sort
 * by key.  Each segment is one rank's mySort (mpi/mpi_lsbsort.cpp:591-597);
 * the reference has no batched form.  With LSB_GRAPH the sequence is a
 * cached hipGraph that is replayed for every batch fitting the library
 * workspace. */
int lsb_sort_segmented(lsb_elem* A, lsb_elem* B, const uint64_t* seg_offsets, uint32_t nseg,
                       const lsb_config* cfg);

/* Device workspace bytes for a caller-provided segmented workspace. */
size_t lsb_segmented_workspace_size(size_t total_n, uint32_t nseg, const lsb_config* cfg);

/* The same batched sort over segments that live in separate buffers:
 * segment i is As[i][0 .. ns[i]), sorted in place, with Bs[i] (at least
 * ns[i] elements, not overlapping As[i]) as its scratch.  Bs[i] may be NULL
 * for segments of at most 8192 elements (sorted inside one workgroup), and
 * Bs itself when every segment is that small.  The pointers are device
 * addresses the caller guarantees (only null, alignment and A/B overlap are
 * checked).  Nothing is copied: callers that hold many separate arrays (the
 * serving loop) batch them without staging copies.  The pointer lists are
 * read before the call returns.  lsb_segmented_workspace_size(sum of ns,
 * nseg, cfg) sizes a caller workspace. */
int lsb_sort_batch(lsb_elem* const* As, lsb_elem* const* Bs, const uint64_t* ns, uint32_t nseg,
                   const lsb_config* cfg);

/* Frees what the library keeps for sorts on `stream` (a hipStream_t; NULL
 * = the default stream) of the current device: the workspaces it allocated
 * for lsb_sort / lsb_sort_segmented / lsb_sort_batch calls made without a
 * caller workspace, and the graphs captured against them; then returns the
 * unused memory of the library's pool to the device.  Waits for the work
 * queued on `stream` first.  Call it before destroying a stream that sorts
 * ran on, with no other thread sorting on that stream; later sorts on it
 * allocate again.  No counterpart in the reference, whose temporaries are
 * per-call vectors (mpi/mpi_lsbsort.cpp:523-528, 550-555). */
int lsb_release(void* stream);

/* ---------------------------------------------------------------- serving */

/* A sort service over jobs resident in HBM.  Jobs are admitted in arrival
 * order against a memory budget that charges each job its doubled buffers
 * (A + B = 32 bytes per element); small jobs admitted while the device is
 * busy are coalesced into one segmented launch (replayed from a hipGraph),
 * larger ones run as plain sorts; launches spread over several streams so
 * their work overlaps.  Each job is sorted in place exactly as lsb_sort
 * would sort it alone.  No counterpart in the reference (one array per run). */
typedef struct lsb_server lsb_server;

typedef struct lsb_server_config {
  size_t memory_budget;      /* bytes; 0 = 200e9 */
  uint32_t max_pending;      /* submitted, not yet admitted; 0 = 65536 */
  uint32_t batch_max_elems;  /* jobs of at most this many elements are coalesced; 0 = 2^24 */
  uint64_t batch_max_total;  /* elements per coalesced launch; each stream owns a scratch slab
                                of this many elements (16 B each); 0 = 2^28 */
  uint32_t streams;          /* device streams; 0 = 8 */
  uint32_t use_graphs;       /* 1 = coalesced launches replay captured hipGraphs */
  uint32_t batch_window_us;  /* longest wait before an open batch is launched; 0 = 200 */
  uint32_t reserved;
} lsb_server_config;

typedef struct lsb_server_stats {
  uint64_t jobs_done;
  uint64_t elems_done;
  uint64_t batches;          /* coalesced launches */
  uint64_t batched_jobs;     /* jobs that ran inside them */
  uint64_t plain_jobs;       /* jobs that ran as plain sorts */
  uint64_t memory_in_use;    /* budget bytes currently charged */
  double p50_us, p99_us, max_us, mean_us;  /* completion - submission */
  double span_s;             /* last completion - first submission */
  uint64_t pool_plain_jobs;  /* plain jobs larger than a scratch slab (scratch from the pool) */
  uint64_t admission_waits;  /* jobs that waited at the head of the queue for budget memory */
} lsb_server_stats;

void lsb_server_config_init(lsb_server_config* cfg);
int lsb_server_create(const lsb_server_config* cfg, lsb_server** out);
/* Finishes every submitted job, then frees the server. */
int lsb_server_destroy(lsb_server* srv);
/* Queues a sort of A[0..n) (device memory, sorted in place by the low
 * key_bits of key; 0 = 64) in the key order key_flags (0, LSB_KEYS_SIGNED or
 * LSB_KEYS_FLOAT).  A must stay valid until the job completes.  LSB_EBUSY
 * when max_pending jobs are waiting; LSB_ENOMEM when the job alone exceeds
 * the budget. */
int lsb_server_submit(lsb_server* srv, lsb_elem* A, size_t n, uint32_t key_bits,
                      uint32_t key_flags, uint64_t* job_id);
/* Queues count jobs at once, in order, as count lsb_server_submit calls
 * would (job i: As[i], ns[i], key_bits[i], key_flags[i]; its id in
 * job_ids[i]) but under one lock, so the scheduler sees them together.  All
 * or nothing: on any error no job is queued. */
int lsb_server_submit_batch(lsb_server* srv, lsb_elem* const* As, const size_t* ns,
                            const uint32_t* key_bits, const uint32_t* key_flags, size_t count,
                            uint64_t* job_ids);
/* 1 = done, 0 = not yet, negative = the job failed with that code.  A
 * finished job stays known until 2^18 later jobs have finished or the
 * statistics are reset; after that LSB_ENOTFOUND. */
int lsb_server_poll(lsb_server* srv, uint64_t job_id);
/* Waits up to timeout_s (< 0: forever); LSB_EBUSY on timeout. */
int lsb_server_wait(lsb_server* srv, uint64_t job_id, double timeout_s);
/* Waits until every submitted job has completed. */
int lsb_server_drain(lsb_server* srv);
int lsb_server_latency(lsb_server* srv, uint64_t job_id, double* us);
/* Counts since the last reset; the latency figures cover the most recent
 * 2^18 completed jobs. */
int lsb_server_stats_get(lsb_server* srv, lsb_server_stats* out);
/* Clears the statistics and forgets completed jobs. */
int lsb_server_reset_stats(lsb_server* srv);

/* Kernel timings of the last lsb_sort call made with LSB_PROFILE on this
 * thread, from HIP events recorded on its stream between its launches.  No
 * counterpart in the reference, which times only the whole sort with
 * steady_clock between barriers (mpi/mpi_lsbsort.cpp:731-745).
 * lsb_profile_read covers a call that ran the LSD form, lsb_msd_profile_read
 * one that ran the MSD form; each returns LSB_EINVAL after the other form.
 * Both read the call's workspace: read before the next sort on the stream. */
typedef struct lsb_profile {
  uint32_t passes;         /* digit passes planned */
  uint32_t executed_mask;  /* bit p set when pass p ran (not single-bucket) */
  uint32_t digit_bits;
  uint32_t tile;
  float histogram_ms;      /* counter reset + all-digit histogram */
  float scan_ms;           /* scan + plan */
  float copy_ms;           /* B -> A copy-back (near 0 when not needed) */
  float total_ms;          /* first event to last */
  float pass_ms[16];       /* scatter pass p */
} lsb_profile;

/* Waits for the profiled call's events and fills *out. */
int lsb_profile_read(lsb_profile* out);

/* The MSD form, kernel by kernel: the set-up, wide step 0 (ranges of
 * 2^21..2^23 elements partitioned on a 10-bit digit), then level k (8-bit
 * digit; the last level only copies) followed by wide step k+1 (level k's
 * buckets of 2^21..2^23 elements) for k + 1 < wide_steps, and the in-LDS
 * finish of every run.  A step is a histogram, a plan and a scatter launch.
 * The element counts are what each scatter moved (partitioned or copied)
 * and what the runs of the finish held: a scatter's and the finish's
 * algorithmic traffic is 32 B per element (16 read, 16 written), a
 * histogram's 8 B per element of its ranges. */
#define LSB_MSD_MAX_LEVELS 9
#define LSB_MSD_WIDE_STEPS 3
typedef struct lsb_msd_profile {
  uint32_t levels;                           /* levels launched */
  uint32_t wide_steps;                       /* wide steps launched */
  uint32_t runs;                             /* runs the finish sorted */
  uint32_t wide_elems[LSB_MSD_WIDE_STEPS];   /* elements wide step s moved */
  uint32_t level_elems[LSB_MSD_MAX_LEVELS];  /* elements level k moved */
  uint32_t finish_elems;                     /* elements of the runs */
  float setup_ms;                            /* small segments + level-0 set-up */
  float wide_ms[LSB_MSD_WIDE_STEPS][3];      /* wide step s: histogram, plan, scatter */
  float level_ms[LSB_MSD_MAX_LEVELS][3];     /* level k: histogram, plan, scatter */
  float finish_ms;
  float total_ms;                            /* first event to last */
} lsb_msd_profile;
int lsb_msd_profile_read(lsb_msd_profile* out);

/* Debug entry for the all-digit histogram kernel alone: writes
 * counts[p * 2^d + b] (uint32, device memory, P = ceil(key_bits/d) rows) for
 * the digits of A[0..n).  The reference's count loop is
 * mpi/mpi_lsbsort.cpp:237-241, run once per digit. */
int lsb_debug_histogram(const lsb_elem* A, size_t n, const lsb_config* cfg,
                        uint32_t* counts);

/* Debug entry for the memory floor of one scatter pass: moves src[0..n) to
 * dst with the pass's access pattern and none of its sorting -- tiles of
 * 8192 elements read in order, each written as `runs` runs of consecutive
 * elements, run r of tile t at dst[r * (n / runs) + t * (8192 / runs)]
 * (runs = 256: the layout one pass over uniform keys writes; runs = 1: a
 * plain copy).  n must be a multiple of 8192 and runs a power of two of at
 * most 8192 (LSB_EINVAL otherwise).  No counterpart in the reference. */
int lsb_debug_scatter_floor(const lsb_elem* src, lsb_elem* dst, size_t n, uint32_t runs,
                            void* stream);

/* ------------------------------------------------------------ workloads */

/* Synthetic inputs, generated on the device.  UNIFORM reproduces the
 * reference's keygen (mpi/mpi_lsbsort.cpp:676-704): the array is block
 * distributed over `ranks` virtual ranks of ceil(n/ranks) elements
 * (divCeil, 146-171); element g takes output number (g mod per) of
 * pcg64(seed_base + g/per) as its key (masked to key_bits) and
 * val = val_base + g. */
#define LSB_GEN_UNIFORM 0u
#define LSB_GEN_SORTED 1u   /* key = i */
#define LSB_GEN_REVERSE 2u  /* key = n-1-i */
#define LSB_GEN_ZIPF 3u     /* key = table[rank], rank ~ Zipf(s) over m keys by inverse CDF of
                               the element's pcg64 draw (53-bit uniform); table[k] = output k
                               of pcg64(zipf_table_seed) */

typedef struct lsb_gen_spec {
  uint32_t kind;
  uint32_t ranks;           /* virtual ranks (>= 1) */
  uint32_t key_bits;        /* UNIFORM: keys masked to the low key_bits bits; 0 = 64 */
  uint32_t zipf_m;          /* ZIPF: number of distinct keys */
  double zipf_s;            /* ZIPF: exponent */
  uint64_t zipf_table_seed; /* ZIPF: seed of the key table */
  uint64_t seed_base;       /* rank r uses pcg64(seed_base + r) */
  uint64_t val_base;        /* val = val_base + global index */
} lsb_gen_spec;

int lsb_generate(lsb_elem* A, size_t n, const lsb_gen_spec* spec, void* stream);

/* O(n) on-device proof that out[0..n) is the stable sort by (masked) key of
 * an input whose vals were val_base + i and whose key at index i is
 * in_keys[i] (device): vals form a permutation, every key matches its source,
 * keys are non-decreasing and equal keys keep index order.  Equivalent to the
 * reference's full stable_sort comparison (mpi/mpi_lsbsort.cpp:755-784).
 * Synchronous.  *bad_count = failing positions, *first_bad = lowest one
 * (or -1). */
int lsb_verify(const lsb_elem* out, const uint64_t* in_keys, size_t n,
               uint64_t val_base, uint32_t key_bits, uint32_t key_flags,
               uint64_t* bad_count, int64_t* first_bad, void* stream);

/* Copies the keys of A[0..n) into keys[0..n) (device), stream-ordered. */
int lsb_extract_keys(const lsb_elem* A, size_t n, uint64_t* keys, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LSB_H */
