/*
 * gpuseed.h - C ABI of the MI355X FM-index seeder (libgpuseed.so).
 *
 * Plain C types only: pointers, sizes, fixed-width integers.  The library owns
 * all device memory; a caller owns a gs_result until gs_result_free().
 * No function calls exit(); every failure is a negative return code that the
 * seed_gen command line maps onto the reference's exit statuses.
 */
#ifndef GPUSEED_H
#define GPUSEED_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define GS_API __attribute__((visibility("default")))
#else
#define GS_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---------------------------------------------------- */
#define GS_OK 0
#define GS_EINVAL -1     /* bad argument                                        */
#define GS_ENOMEM -2     /* device or host allocation failed                    */
#define GS_EHIP -3       /* a HIP runtime call failed                           */
#define GS_EIO -4        /* index file missing or unreadable                    */
#define GS_EFORMAT -5    /* .sa/.bwt inconsistent (primary or seq_len differ)  */
#define GS_ENODEV -6     /* no usable gfx950 device                             */
#define GS_ERANGE -7     /* input beyond a documented limit                     */

/* ---- random-access roofline (measurement support) ---------------------- */
typedef struct {
  double seconds;     /* best timed launch                                    */
  double lines_per_s; /* random line reads served per second                  */
  double bytes_per_s; /* lines_per_s * line_bytes                             */
} gs_roofline_result;

/* Gather microbenchmark over a buffer of `buffer_bytes` filled with random data.
 * mode 0: each lane keeps `inflight` (1, 2 or 4) independent line reads in
 *         flight per round, `iters` reads in total (throughput ceiling).
 * mode 1: each lane follows a dependent chain of `iters` reads (the next index
 *         depends on the data just read).
 * grid = number of 256-lane workgroups.  line_bytes in {16, 32, 64, 128}.
 * No counterpart in the reference (it publishes no roofline). */
GS_API int gs_roofline_random(uint64_t buffer_bytes, int line_bytes, int mode, int inflight, int grid,
                       int iters, int reps, gs_roofline_result* out);

/* Streaming device-to-device copy of `bytes`, best of `reps`; used to calibrate
 * hardware byte counters against a kernel of known traffic.  No counterpart in
 * the reference. */
GS_API int gs_stream_copy_bytes(uint64_t bytes, int reps, double* seconds_out);

/* ---- synthetic workloads ----------------------------------------------
 * The reference ships no data; these deterministic generators make the
 * references and reads of the benchmark configurations.  Output depends only
 * on the parameters and seed, never on n_threads (0 = pick automatically).
 * No counterpart in the reference. */
typedef struct {
  uint64_t length; /* bases                                                   */
  uint64_t seed;
  double gc; /* P(C)+P(G); A and T share the rest equally                     */
  /* interspersed repeat families; copies overwrite random positions */
  uint32_t n_families, family_len_min, family_len_max;
  uint32_t copies_per_family;   /* if > 0: exactly this many copies per family */
  double interspersed_fraction; /* else: place copies until this fraction     */
  double family_div_min, family_div_max; /* per-copy substitution rate       */
  /* exact tandem repeats */
  double tandem_fraction;
  uint32_t tandem_unit_min, tandem_unit_max, tandem_copies_min, tandem_copies_max;
  /* segmental duplications copied from the genome itself */
  double segdup_fraction;
  uint32_t segdup_len_min, segdup_len_max;
  double segdup_div_min, segdup_div_max;
} gs_synth_ref_params;

/* Fills codes[0..p->length) with 0..3 (A,C,G,T). */
GS_API int gs_synth_reference(const gs_synth_ref_params* p, uint8_t* codes, int n_threads);

typedef struct {
  double weight;
  uint32_t len_min, len_max; /* uniform length                               */
  double sub_rate;           /* substitutions per base                       */
  double indel_rate;         /* indel events per base, 1-3 bp, ins:del 1:1   */
} gs_read_class;

typedef struct {
  uint64_t n_reads;
  uint64_t seed;
  uint32_t n_classes; /* 1..4 */
  gs_read_class classes[4];
  double n_read_fraction; /* fraction of reads with one base replaced by 'N'  */
  int both_strands;       /* 1: strand drawn per read; 0: forward only        */
} gs_synth_read_params;

/* Read lengths only (cheap), so a caller can size and lay out buffers. */
GS_API int gs_synth_read_lengths(const gs_synth_read_params* p, uint32_t* lens);
/* ASCII reads written at out+offsets[i]; origin_pos/origin_strand optional. */
GS_API int gs_synth_reads(const uint8_t* ref_codes, uint64_t ref_len, const gs_synth_read_params* p,
                   const uint64_t* offsets, char* out, uint32_t* origin_pos, uint8_t* origin_strand,
                   int n_threads);

/* Parameters of benchmark configuration 1..5 (reference genome and reads). */
GS_API int gs_synth_preset(int config, gs_synth_ref_params* ref, gs_synth_read_params* reads);

/* ---- index construction (replaces build_index.sh + the bwa fork) ---------
 * Writes <prefix>.bwt and <prefix>.sa in the layout seed_gen reads
 * (seed_gen.cu:2685-2736) for the text codes[0..n) (values 0..3), with the
 * suffix array sampled every sa_intv rows (build_index.sh:24-63, -r RATIO).
 * CPU SA-IS; n < 2^32 - 2. */
GS_API int gs_build_index_host(const uint8_t* codes, uint64_t n, uint32_t sa_intv, const char* prefix);

/* ---- upstream bwa 0.7 files --------------------------------------------
 * gs_index_open (and gs_index_convert) also read the .bwt/.sa that upstream
 * bwa 0.7's `bwa index` writes: 64-bit integers and 128-base occurrence
 * blocks instead of the fork's 32-bit integers and 64-base blocks
 * (seed_gen.cu:2685-2736), recognised by header and file size.  Its text is
 * usually the reference followed by its reverse complement; any text of fewer
 * than 2^32 - 2 bases loads (longer ones give GS_ERANGE), and seeds are then
 * reported against that text.
 * No counterpart in the reference, which reads only the fork's files. */

/* Writes <dst_prefix>.bwt/.sa in the fork's layout from either layout at
 * <src_prefix> (CPU only).  GS_EIO, GS_EFORMAT as gs_index_open. */
GS_API int gs_index_convert(const char* src_prefix, const char* dst_prefix);

/* bwa's reference files <prefix>.pac/.ann/.amb: the forward reference as codes
 * 0..3 (runs of N hold the random bases bwa put there; .amb lists them), the
 * sequences' names and offsets in the concatenated text.  The codes can be
 * indexed with gs_index_build / gs_build_index_host, which gives the index the
 * fork's build_index.sh makes for the same FASTA; seed positions are offsets
 * into that concatenation, and gs_bwa_ref_locate maps them to a sequence.
 * CPU only.  No counterpart in the reference (which prints raw positions). */
typedef struct gs_bwa_ref gs_bwa_ref;
GS_API int gs_bwa_ref_load(const char* prefix, gs_bwa_ref** out);
GS_API void gs_bwa_ref_free(gs_bwa_ref* ref);
GS_API uint64_t gs_bwa_ref_length(const gs_bwa_ref* ref);
GS_API const uint8_t* gs_bwa_ref_codes(const gs_bwa_ref* ref);
GS_API int gs_bwa_ref_n_contigs(const gs_bwa_ref* ref);
GS_API int gs_bwa_ref_contig(const gs_bwa_ref* ref, int i, const char** name, uint64_t* offset, uint64_t* len);
GS_API int gs_bwa_ref_n_holes(const gs_bwa_ref* ref);
GS_API int gs_bwa_ref_hole(const gs_bwa_ref* ref, int i, uint64_t* offset, uint64_t* len, char* c);
/* pos in [0, length): the sequence holding it and the offset inside it;
 * GS_ERANGE beyond the end. */
GS_API int gs_bwa_ref_locate(const gs_bwa_ref* ref, uint64_t pos, int* contig, uint64_t* offset);

/* ---- the device-resident index -----------------------------------------
 * gs_index_open replaces bwt_restore_bwt + bwt_restore_sa + the upload and the
 * 13-mer table of seed_gen (seed_gen.cu:2685-2736, 2840-2895).  It reads
 * <prefix>.bwt and <prefix>.sa, checks them like the reference (GS_EIO for a
 * missing file, GS_EFORMAT for primary/seq_len disagreement, seed_gen.cu:
 * 2692-2707) and additionally rejects occurrence counts that disagree with the
 * BWT, then repacks the occurrence blocks into the device layout. */
typedef struct gs_index gs_index;

typedef struct {
  uint32_t seq_len, primary, sa_intv;
  uint32_t L2[5];
  uint64_t n_blocks;     /* 96-base occurrence blocks on the device          */
  uint64_t device_bytes; /* HBM held by the index                            */
  double t_read, t_upload, t_repack, t_table; /* seconds                      */
  double t_sa, t_write; /* gs_index_build: suffix sorting, file writing      */
  int sa_rounds;        /* gs_index_build: prefix-doubling rounds            */
  int kmer_len;         /* longest interval table on the device (13..16)     */
  double t_ext;         /* gs_index_set_kmer_len: table build, seconds       */
  int direct;           /* 1: direct-extension structures on the device      */
  double t_direct;      /* gs_index_set_direct: build, seconds               */
  int pairs;            /* 1: pair table on the device                       */
  double t_pairs;       /* gs_index_set_pairs: build, seconds                */
} gs_index_info;

GS_API int gs_index_open(const char* prefix, int device, gs_index** out);
/* Drops the caller's handle.  Batches and servers made on the index keep it
 * alive: its device memory is freed with the last of them. */
GS_API int gs_index_close(gs_index* idx);
GS_API int gs_index_get_info(const gs_index* idx, gs_index_info* info);
/* Copies device structures back for inspection (tests): occurrence entries
 * (4 u32 each: count, 96-bit bitmap) and 13-mer table entries (2 u32: l, u). */
GS_API int gs_index_read_occ(const gs_index* idx, uint64_t first, uint64_t count, uint32_t* out4);
GS_API int gs_index_read_kmer(const gs_index* idx, uint64_t first, uint64_t count, uint32_t* out2);

/* Adds (len 14..16) or drops (len 13) an extended interval table: the SA
 * interval of every pattern of len bases and which of its 64 three-base left
 * extensions occur, 16 * 4^len bytes of HBM (4.3 GB, 17.2 GB, 68.7 GB).
 * Seeding then answers a task's first len bases with one lookup where the
 * reference takes the 13-mer table plus len - 13 backward steps
 * (seed_gen.cu:2132-2174).  A task whose clean window is absent (with
 * k >= len), or whose clean (len + 3)-base window is absent (with
 * k >= len + 3), cannot reach k bases and ends at once; otherwise it falls
 * back to those steps when the pattern does not occur or its window holds a
 * code above 3.  The output is unchanged.  No counterpart in the reference. */
GS_API int gs_index_set_kmer_len(gs_index* idx, int len);
/* gs_index_set_kmer_len, gs_index_set_direct and gs_index_set_pairs change
 * structures that every batch and server of the index reads: they return
 * GS_EINVAL while a server made on the index is alive, and wait for all
 * device work of the process before freeing anything; the caller must not
 * run a batch of the index on another thread during the call. */
/* Adds (on != 0) or drops the direct-extension structures: the full suffix
 * array, its inverse and the text at 2 bits per base, derived on the device
 * from the index itself (8 * (seq_len + 1) + seq_len / 4 bytes of HBM: 25.6 GB
 * at 3.1 Gbp).  The batched kernels then finish a task whose interval has kept
 * a single row for gs_opts.direct_min backward steps by reading that row's
 * text position, comparing the read with the text 64 bases per 16-byte load
 * and reading the row where the match ends, in place of one backward step
 * per base.  The output is unchanged.  No counterpart in the reference. */
GS_API int gs_index_set_direct(gs_index* idx, int on);
/* Adds (on != 0) or drops the pair table: one byte per 18-base pattern
 * ("core"), bits 0-3 saying which bases follow it in the text, bits 4-7
 * which precede it (4^18 bytes = 68.7 GB of HBM, at any text length).  With
 * the 16-mer extended table and k >= 19, once a task of a read strand has
 * found its clean 19-base window absent, the next task reads the byte of its
 * window's core instead, and with 22 clean codes also the byte of the core
 * two bases to its left, in the same load: one operation answers whether the
 * windows of up to four consecutive task ends occur, so on a strand that does
 * not match the reference up to four tasks end per lookup instead of one.
 * The output is unchanged.  Built from
 * the text of the direct-extension structures, or from a text derived from
 * the index by the same walk.  No counterpart in the reference. */
GS_API int gs_index_set_pairs(gs_index* idx, int on);
/* Bytes [first, first+count) of the pair table (tests). */
GS_API int gs_index_read_pairs(const gs_index* idx, uint64_t first, uint64_t count, uint8_t* out);
/* Text positions of n suffix-array rows as the reference's walk reports
 * them, row primary+1 behaviour included (locate_seeds_gpu with
 * bwt_inv_psi_gpu, seed_gen.cu:455-484, 668-684); rows past seq_len and row
 * 0 give 0xFFFFFFFF.  From the full suffix array when the index has the
 * direct-extension structures, else by the sampled-suffix-array walk. */
GS_API int gs_index_locate(gs_index* idx, const uint32_t* rows, uint64_t n, uint32_t* out);
/* Rows [first, first+count) of the full suffix array and positions
 * [first, first+count) of its inverse and of the text (one code 0..3 per
 * byte); any output may be NULL (tests). */
GS_API int gs_index_read_direct(const gs_index* idx, uint32_t first, uint32_t count, uint32_t* sa_out,
                                uint32_t* isa_out, uint8_t* text_out);
/* Entries [first, first+count) of the extended table (tests): four words
 * per entry, l, u and the low and high words of the presence mask (bit
 * c1 + 4*c2 + 16*c3 set iff the pattern extended to the left by c1, then c2,
 * then c3 occurs). */
GS_API int gs_index_read_kext(const gs_index* idx, uint64_t first, uint64_t count, uint32_t* out4);
/* Entries [first, first+count) of the tandem table (tests): four words per
 * entry, l, u, m, 0.  Entry key * 237 + (n - 20), n in [20, 256], of unit
 * key (p bases for key in [(4^p - 4)/3, (4^(p+1) - 4)/3), p = 1..6, code
 * key - (4^p - 4)/3 = sum of code_j * 4^j with the unit's last base as j = 0)
 * describes the string of length n ending with the unit repeated: m is the
 * longest length <= n whose string occurs in the text and [l, u] its
 * interval.  Built with every index.  No counterpart in the reference. */
GS_API int gs_index_read_tandem(const gs_index* idx, uint64_t first, uint64_t count, uint32_t* out4);

/* Builds the index of codes[0..n) (values 0..3) on the device: suffix array
 * by prefix doubling, then the same device structures gs_index_open makes.
 * With write_prefix, also writes <write_prefix>.bwt and .sa in the reference
 * layout (byte-identical to gs_build_index_host).  Replaces the bwa fork that
 * build_index.sh drives (build_index.sh:54-75).  n < 2^32 - 2. */
GS_API int gs_index_build(const uint8_t* codes, uint64_t n, uint32_t sa_intv, int device,
                          const char* write_prefix, gs_index** out);

/* ---- seeding --------------------------------------------------------------
 * gs_seed replaces one run of seed_gen's per-batch pipeline (seed_gen.cu:
 * 2910-3790: find_seed_intervals_gpu, the compaction and filters, the
 * segmented sort, seeds_to_threads(_mem) and locate_seeds_gpu).  Seeds come out
 * per read in the reference's order: merged by qbeg with forward seeds first
 * on ties, positions of a seed in suffix-array row order. */
typedef struct {
  int k;            /* -k: minimum seed length, >= 14 (seed_gen.cu:2821)       */
  int smem;         /* -s: SMEM (1) or MEM (0) mode (2832)                     */
  uint32_t max_occ; /* 0: every row located (the reference); else seeds whose
                       interval is larger get no positions                      */
  uint32_t batch_bases; /* MEM look-ahead stays inside batches closed at this
                       many bases (seed_gen.cu:2814, 2955); 0 = 1,000,000     */
  int kernel;       /* 0: plain kernel, one stream per lane (the drop-in);
                       1: batched persistent kernel, 8-bit size records;
                       2: batched persistent kernel, 16-bit size records;
                       3: batched persistent kernel, 4-bit size records (a
                          smaller LDS slot, so more waves per CU)             */
  uint32_t seg_tasks; /* batched kernels: reads longer than 160 bases are cut
                       into work items of this many task ends, 1..112
                       (0 = 64)                                               */
  uint16_t wg;        /* batched kernels: workgroup size 64, 128 or 256 (0 = 128) */
  uint16_t walk_rows; /* suffix-array walk: rows per thread 1, 2 or 4 (0 = 1) */
  uint16_t split_tasks; /* batched kernels: in the tail of a launch an item
                       that has used split_work memory operations hands half
                       of its task ends to an idle lane while at least
                       2 * split_tasks remain, 1..64; 0 = 8, and only in
                       batches of at most 4 work items per lane of the
                       persistent grid; 0xFFFF: no splitting                 */
  uint16_t split_work;  /* in units of 64 memory operations, 1..0xFFFE (0 = 32) */
  uint16_t direct_min;  /* batched kernels on an index with gs_index_set_direct:
                       a task reads the text once its interval has kept one
                       row for this many backward steps, 1..64 (0 = 4);
                       0xFFFF: never                                         */
  uint16_t tandem;     /* batched kernels: tandem jumps for repeat periods up
                       to this, 1..6 (0 = 6); 0xFFFF: off.  Inside a stretch
                       read[R, E) of the read with period p, read[y, e) for
                       y >= R is the last p bases of read[.., e) repeated,
                       so its interval depends only on that unit and on
                       e - y: a task end e of the stretch reads the index's
                       tandem table (about 21 MB, built with the index) for
                       length min(e - R, 256) instead of its lookup and the
                       steps down to R, and goes on from there (or ends,
                       where the table says the repeat stops occurring).
                       The output depends on none of the tuning fields
                       (kernel, seg_tasks, wg, walk_rows, split_*,
                       direct_min, tandem)                                   */
} gs_opts;

typedef struct {
  uint32_t qbeg, qend; /* bit 31 set on both for the reverse strand (2402)   */
  uint32_t l, u;       /* suffix-array interval rows                           */
  uint32_t n_pos;      /* positions reported for this seed                     */
  uint32_t pad;
} gs_seed_rec;

typedef struct {
  uint64_t n_reads, n_seeds, n_pos;
  uint32_t* read_nseeds;   /* [n_reads]                                       */
  uint32_t* read_npos;     /* [n_reads]                                       */
  uint64_t* read_seed_off; /* [n_reads + 1]                                   */
  uint64_t* read_pos_off;  /* [n_reads + 1]                                   */
  gs_seed_rec* seeds;      /* [n_seeds]                                       */
  uint32_t* pos;           /* [n_pos] reference positions                     */
  /* device time of each phase, seconds (HIP events on the index's stream) */
  double t_h2d, t_seed, t_post, t_locate, t_d2h;
} gs_result;

/* The device buffers of a call stay with the index for the next one (they
 * only grow; gs_index_close frees them). */
GS_API int gs_seed(gs_index* idx, const char* bases, const uint64_t* offsets, const uint32_t* lens,
                   uint64_t n_reads, const gs_opts* opts, gs_result* out);
GS_API void gs_result_free(gs_result* res);

/* gs_seed for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it): no hipMalloc, hipFree or
 * other allocation, on the device or the host.  The device buffers of the
 * call (reads, task maps, candidates, counts, seed records, rows) are slices
 * of a fixed 256 MiB array that each GPU the library runs on holds until the
 * process ends; the results go into the caller's arrays below.  A call whose
 * buffers do not fit the array returns GS_ERANGE (about 7 million bases of
 * 150 bp reads per call; candidates take 32 bytes per base); so does a call
 * whose seeds or positions exceed seeds_cap or pos_cap, with n_seeds and
 * n_pos set to what it needs.  Output equals gs_seed's.  The persistent
 * kernel runs when every read fits its 160-base window (opts->kernel 1..3),
 * else the plain kernel.  Calls on one GPU share the array: calls that
 * overlap on one GPU race.  Runs on the index's stream.  No counterpart in
 * the reference. */
typedef struct {
  uint32_t* read_nseeds;   /* [n_reads]                                       */
  uint32_t* read_npos;     /* [n_reads]                                       */
  uint64_t* read_seed_off; /* [n_reads + 1]                                   */
  uint64_t* read_pos_off;  /* [n_reads + 1]                                   */
  gs_seed_rec* seeds;      /* [seeds_cap]                                     */
  uint32_t* pos;           /* [pos_cap]                                       */
  uint64_t seeds_cap, pos_cap;
  uint64_t n_seeds, n_pos; /* set by the call                                 */
} gs_result_buffers;

GS_API int gs_seed_noalloc(gs_index* idx, const char* bases, const uint64_t* offsets, const uint32_t* lens,
                           uint64_t n_reads, const gs_opts* opts, gs_result_buffers* out);

/* Device-resident batches: gs_seed is gs_batch_create + gs_batch_run +
 * gs_batch_fetch.  gs_batch_create uploads the reads once; gs_batch_run runs
 * the whole device pipeline and leaves the results in HBM (it may be called
 * again, e.g. for timing); gs_batch_fetch copies them to a gs_result.  Each
 * batch has a HIP stream of its own, so batches of one index may run from
 * different threads at once; one batch must not be used by two threads at
 * once.  A batch keeps its index alive until gs_batch_destroy.  No
 * counterpart in the reference, whose batches are host round trips
 * (seed_gen.cu:3108-3601). */
typedef struct gs_batch gs_batch;

typedef struct {
  double t_h2d;    /* upload at gs_batch_create                              */
  double t_seed;   /* backward-extension kernel                              */
  double t_count;  /* per-read counts and scans                              */
  double t_write;  /* seed records and suffix-array rows                      */
  double t_locate; /* suffix-array walk                                       */
  double t_total;  /* first kernel start to last kernel end                   */
  uint64_t n_reads, n_seeds, n_pos;
  uint64_t n_table_reads, n_occ_reads; /* memory operations of the batched
                      extension kernel (0 for the plain kernel)                */
  uint64_t n_direct_reads; /* its suffix-array, text and inverse reads       */
  uint64_t n_tandem_reads; /* its tandem-table entries read (gs_opts.tandem) */
  double t_kernel; /* the extension kernel's launch alone (seed_persistent, or
                      the plain kernel): t_seed less code packing, the work
                      queue's set-up and the join of split streams; 0 from
                      gs_batch_replay, whose graph holds no timing events)  */
} gs_batch_stats;

GS_API int gs_batch_create(gs_index* idx, const char* bases, const uint64_t* offsets,
                           const uint32_t* lens, uint64_t n_reads, const gs_opts* opts,
                           gs_batch** out);
GS_API int gs_batch_run(gs_batch* batch, int kernel, gs_batch_stats* stats); /* kernel -1: as created */
GS_API int gs_batch_fetch(gs_batch* batch, gs_result* out);
GS_API int gs_batch_destroy(gs_batch* batch);

/* The whole pipeline as one HIP graph: upload of the reads (from pinned host
 * copies of the batch's inputs made here), extension, counts and scans,
 * records, suffix-array walk and the download of every result array into
 * pinned host memory.  Output sizes come from the batch's previous
 * gs_batch_run, so the graph has no host synchronisation inside.
 * gs_batch_replay launches it and waits; gs_batch_fetch then returns the
 * downloaded results.  Replays write the same results as gs_batch_run.
 * kernel -1: as created.  No counterpart in the reference, which
 * synchronises after every phase. */
GS_API int gs_batch_capture(gs_batch* batch, int kernel);
GS_API int gs_batch_replay(gs_batch* batch, gs_batch_stats* stats);

/* Autotuner: times the batched pipeline on the caller's reads (a sample of
 * the workload) over its launch parameters: seg_tasks (if some reads are
 * longer than 160 bases), size-record width x workgroup size, and suffix-array
 * walk rows, one after another, each candidate the best of `reps` runs
 * (0 = 3).  Returns base with the chosen tuning fields.  Every candidate must
 * give the same seed and position counts (the output does not depend on the
 * tuning fields).  No counterpart in the reference. */
typedef struct {
  gs_opts opts;           /* base with kernel, seg_tasks, wg, walk_rows, direct_min chosen */
  double seconds;         /* pipeline time with opts                           */
  double default_seconds; /* pipeline time with base's fields (0 -> defaults)  */
  int n_tried;
  int pad;
} gs_tune_result;

GS_API int gs_autotune(gs_index* idx, const char* bases, const uint64_t* offsets, const uint32_t* lens,
                       uint64_t n_reads, const gs_opts* base, int reps, gs_tune_result* out);

/* ---- serving ---------------------------------------------------------------
 * A server keeps the index resident and seeds requests submitted from any
 * thread with continuous batching: a planner thread packs the unscheduled
 * reads of every waiting request into device rounds (each request gets a fair
 * share of a round, whole 1,000,000-base reference batches at a time, then the
 * round is filled in arrival order), bounded by round_bases and by a share of
 * budget_bytes of device memory for read and seed buffers (the budget over
 * workers + 1 rounds); device threads run rounds as one pipeline each (one
 * persistent-kernel launch over all its requests), three rounds on the GPU at
 * once by default, while the planner prepares the next.  Requests arriving
 * during a round join the next one.  A request's results are exactly those of gs_seed
 * on the request alone.  The caller's read buffers must stay valid until
 * gs_server_wait returns for that request.  No counterpart in the reference. */
typedef struct gs_server gs_server;

typedef struct {
  gs_opts opts;          /* seeding options of every request; kernel 0 means 3
                            unless exclusive, seg_tasks 0 means 32            */
  uint64_t budget_bytes; /* device memory of the rounds in flight (0 = 64 GiB) */
  uint64_t round_bases;  /* bases per round at most (0 = 16,000,000)         */
  int exclusive;         /* 1: a round holds one whole request, in arrival order
                            (the per-request baseline)                        */
  int workers;           /* device threads, each with its own stream and round
                            buffers (0 = 3, or 1 when exclusive; at most 8)   */
  int round_growth;      /* rounds grow to backlog / workers bases, at most
                            round_growth x round_bases, at most 64 (0 = 4, or
                            1 when exclusive; 1 = fixed size): more throughput
                            under load, longer waits for small requests       */
  int reserved;
} gs_server_opts;

typedef struct {
  double t_submit, t_start, t_done; /* seconds on the server clock; t_start =
                                       start of the request's first round    */
  uint64_t n_reads, n_seeds, n_pos;
  uint32_t n_rounds, pad;
} gs_request_info;

typedef struct {
  uint64_t n_requests, n_rounds, n_reads, n_bases;
  double busy_seconds;       /* wall time of device rounds (upload to download) */
  double device_seconds;     /* pipeline time of the rounds (HIP events)       */
  double t_load, t_run, t_fetch, t_split; /* wall time of the round stages, summed */
  uint64_t peak_round_bytes; /* largest round's device buffers, outputs excluded */
  double elapsed;            /* server clock now                                */
  double t_plan;             /* planner: choosing and planning rounds, summed  */
  double t_idle;             /* device threads waiting for a planned round, summed */
  double t_seed, t_post, t_locate; /* device time of the pipeline stages, summed */
} gs_server_stats;

GS_API int gs_server_create(gs_index* idx, const gs_server_opts* opts, gs_server** out);
GS_API int gs_server_submit(gs_server* srv, const char* bases, const uint64_t* offsets, const uint32_t* lens,
                            uint64_t n_reads, uint64_t* ticket);
/* Blocks until the request is done; each ticket is waited for once. */
GS_API int gs_server_wait(gs_server* srv, uint64_t ticket, gs_result* out, gs_request_info* info);
GS_API int gs_server_get_stats(gs_server* srv, gs_server_stats* stats);
GS_API double gs_server_clock(gs_server* srv);
/* Finishes every submitted request, then stops the threads.  Returns the
 * first failure of any round (that round's requests failed with it; the
 * others completed), else GS_OK. */
GS_API int gs_server_destroy(gs_server* srv);

/* ---- test support ----------------------------------------------------------
 * No counterpart in the reference. */
/* Work-items per launch of the chunked kernels (table levels, presence bits,
 * direct-structure chains, suffix-array walk), so that a small index runs
 * every chunked loop past its first chunk; 0 restores the default 2^30.
 * Process-wide; set it while no device work runs. */
GS_API int gs_set_launch_chunk(uint64_t items);
/* The round_no-th round a server's device threads take (counted from 0)
 * fails with rc (negative) before its upload; its requests fail with rc and
 * later requests are served normally. */
GS_API int gs_server_inject_failure(gs_server* srv, uint64_t round_no, int rc);

/* The -o text of a result (seed_gen.cu:3625-3648) written to file descriptor
 * fd, reads numbered from first_read_no. */
GS_API int gs_format_text(const gs_result* res, int smem, uint64_t first_read_no, int fd);

#ifdef __cplusplus
}
#endif

#endif /* GPUSEED_H */
