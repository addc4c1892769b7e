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

#ifdef __cplusplus
}
#endif

#endif /* GPUSEED_H */
