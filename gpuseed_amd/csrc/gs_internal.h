// Host-side internals shared by the HIP translation units of libgpuseed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpuseed.h"
#include "gs_device.h"

#define GS_HIP(call)                        \
  do {                                      \
    hipError_t e_ = (call);                 \
    if (e_ != hipSuccess) return GS_EHIP;   \
  } while (0)

struct gs_index {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t primary = 0, seq_len = 0, sa_intv = 0;
  uint32_t L2[5] = {0, 0, 0, 0, 0};
  uint64_t n_blocks = 0;  // 96-base blocks
  uint64_t n_sa = 0;
  uint4* occ = nullptr;
  uint32_t* sa = nullptr;
  uint2* kmer = nullptr;
  uint64_t device_bytes = 0;
  double t_read = 0, t_upload = 0, t_repack = 0, t_table = 0, t_sa = 0, t_write = 0;
  int sa_rounds = 0;
  gs::DevIndex dev() const;
};

namespace gs {

// Parameters of the persistent extension kernel (gs_persist.hip).
struct PersistParams {
  DevIndex X;
  const uint4* codes;        // 4-bit virtual codes, per read: forward chunks then reverse chunks
  const uint64_t* code_off;  // per read, in 16-byte chunks
  const uint32_t* lens;
  const uint64_t* task_off;
  uint32_t n_reads, k, smem, pad;
  uint4* cand;               // same candidate layout as the plain kernel
  uint32_t* nF;
  uint32_t* nR;
  uint32_t* queue;           // work-queue counter
  uint64_t* counters;        // optional: [0] table reads, [1] occ entry reads
};

constexpr uint32_t kPersistLmax = 160;  // longest read held in a persistent lane's LDS slot

int launch_pack_codes(const uint8_t* bases, const uint64_t* offs, const uint32_t* lens, const uint64_t* code_off,
                      uint32_t n_reads, uint32_t lmax, uint4* codes, hipStream_t st);
// variant 0: 8-bit size records, 1: 16-bit size records
int launch_persistent(const PersistParams& P, int variant, hipStream_t st);

}  // namespace gs

// ceil(2^64 / d) for d >= 2 (exact quotient for every 32-bit numerator), 0 for d == 1
uint64_t gs_reciprocal64(uint32_t d);

// Builds the 4^13 interval table on the index's stream (pre_calc_seed_intervals_gpu,
// seed_gen.cu:2648-2675), level by level.
int gs_build_kmer_table(gs_index* X);
