// Index construction on the GPU: suffix array by prefix doubling, then the
// device layout (occ96, sampled SA, 13-mer table) and optionally the
// reference-format .bwt / .sa files.
//
// The reference builds its index offline with a bwa fork driven by
// build_index.sh (build_index.sh:54-75); the fork is not vendored, and bwa's
// CPU construction takes on the order of an hour for a 3.1 Gbp genome.  The
// bytes of the index are fully determined by the text, so this builder writes
// the same files (.bwt with 64-base count blocks, .sa sampled every sa_intv
// rows; seed_gen.cu:52-58, 2685-2736) in seconds and can hand the device
// structures straight to the seeder.
//
// Suffix sorting: every suffix gets the key of its first 21 characters (3 bits
// per character, 0 past the end of the text so that shorter suffixes sort
// first, as with bwa's `$`).  One radix sort orders all keys; groups of equal
// keys are then refined by prefix doubling (Manber & Myers 1993; the ranking
// scheme of Larsson & Sadakane 2007): a suffix's rank is the first row of its
// group, and a group's members are re-sorted by (rank[p], rank[p + H]) with H
// doubling each round.  Only unresolved suffixes take part in later rounds.
// The device-wide radix sorts, scans and compactions come from rocPRIM.
#include <stdio.h>
#include <string.h>
#include <sys/time.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <vector>

#include "gs_internal.h"
#include "host/gs_host.h"

namespace {

double wall() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

inline unsigned grid_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

constexpr int kKeyChars = 21;

__global__ void make_keys(const uint8_t* __restrict__ T, uint64_t n, uint64_t* __restrict__ keys,
                          uint32_t* __restrict__ vals) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i > n) return;
  uint64_t key = 0;
  for (int j = 0; j < kKeyChars; ++j) {
    const uint64_t p = i + j;
    key = (key << 3) | (p < n ? (uint64_t)(T[p] + 1) : 0ull);
  }
  keys[i] = key;
  vals[i] = (uint32_t)i;
}

// first-row-of-group marker: row index at a group's first row, 0 elsewhere
__global__ void head_rows(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ rows, uint64_t m,
                          uint32_t* __restrict__ out) {
  const uint64_t j = blockIdx.x * 256ull + threadIdx.x;
  if (j >= m) return;
  const uint32_t row = rows ? rows[j] : (uint32_t)j;
  out[j] = (j == 0 || keys[j] != keys[j - 1]) ? row : 0u;
}

// rank of each suffix = first row of its group; flag members of groups of 2+
__global__ void apply_groups(const uint32_t* __restrict__ rows, const uint32_t* __restrict__ pos,
                             const uint32_t* __restrict__ grp, uint64_t m, uint32_t* __restrict__ sa,
                             uint32_t* __restrict__ rank, uint8_t* __restrict__ flag) {
  const uint64_t q = blockIdx.x * 256ull + threadIdx.x;
  if (q >= m) return;
  const uint32_t row = rows ? rows[q] : (uint32_t)q;
  const uint32_t p = pos[q], g = grp[q];
  sa[row] = p;
  rank[p] = g;
  flag[q] = (q > 0 && grp[q - 1] == g) || (q + 1 < m && grp[q + 1] == g);
}

// An unresolved suffix shares its first H characters with another suffix, so
// it is at least H long and p + H <= n; the clamp only guards the array.
__global__ void round_keys(const uint32_t* __restrict__ rows, uint64_t u, const uint32_t* __restrict__ sa,
                           const uint32_t* __restrict__ rank, uint64_t H, uint64_t n,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ pos) {
  const uint64_t q = blockIdx.x * 256ull + threadIdx.x;
  if (q >= u) return;
  const uint32_t p = sa[rows[q]];
  const uint64_t p2 = (uint64_t)p + H;
  keys[q] = ((uint64_t)rank[p] << 32) | (p2 <= n ? rank[p2] : 0u);
  pos[q] = p;
}

__global__ void find_primary(const uint32_t* __restrict__ sa, uint64_t m, uint32_t* __restrict__ out) {
  const uint64_t j = blockIdx.x * 256ull + threadIdx.x;
  if (j < m && sa[j] == 0) *out = (uint32_t)j;
}

// bitmaps and per-block counts of the $-removed BWT in 96-base blocks
__global__ void bwt_blocks96(const uint8_t* __restrict__ T, const uint32_t* __restrict__ sa, uint32_t n,
                             uint32_t primary, uint64_t nb, uint4* __restrict__ occ,
                             uint32_t* __restrict__ bcnt) {
  const uint64_t b = blockIdx.x * 256ull + threadIdx.x;
  if (b >= nb) return;
  uint32_t bits[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  uint32_t cnt[4] = {0, 0, 0, 0};
  for (uint32_t q = 0; q < gs::kBlockBases; ++q) {
    const uint64_t i = b * gs::kBlockBases + q;
    if (i >= n) break;
    const uint64_t row = i + (i >= primary ? 1 : 0);
    const uint32_t c = T[sa[row] - 1];
    bits[c][q >> 5] |= 1u << (q & 31);
    ++cnt[c];
  }
  for (uint32_t c = 0; c < 4; ++c) {
    occ[b * 4 + c] = make_uint4(0, bits[c][0], bits[c][1], bits[c][2]);
    bcnt[c * nb + b] = cnt[c];
  }
}

__global__ void set_block_counts(const uint32_t* __restrict__ bcnt_scanned, uint64_t nb, uint4* __restrict__ occ) {
  const uint64_t b = blockIdx.x * 256ull + threadIdx.x;
  if (b >= nb) return;
  for (uint32_t c = 0; c < 4; ++c) occ[b * 4 + c].x = bcnt_scanned[c * nb + b];
}

__global__ void sample_sa(const uint32_t* __restrict__ sa, uint64_t n_sa, uint32_t intv, uint32_t* __restrict__ out) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n_sa) return;
  out[i] = i == 0 ? 0xFFFFFFFFu : sa[i * intv];
}

// reference-layout body: 64-base blocks of 4 counts + 4 MSB-first words
__global__ void ref_blocks64(const uint4* __restrict__ occ, uint32_t n, uint64_t n64,
                             uint32_t* __restrict__ body) {
  const uint64_t j = blockIdx.x * 256ull + threadIdx.x;
  if (j >= n64) return;
  const uint64_t p0 = j * 64;
  // counts before p0 from the 96-base entries
  const uint64_t b = p0 / gs::kBlockBases, r = p0 - b * gs::kBlockBases;
  for (uint32_t c = 0; c < 4; ++c) {
    const uint4 e = occ[b * 4 + c];
    body[j * 8 + c] = r ? gs::rank_in_entry(e, (uint32_t)r - 1) : e.x;
  }
  for (uint32_t w = 0; w < 4; ++w) {
    uint32_t word = 0;
    for (uint32_t q = 0; q < 16; ++q) {
      const uint64_t i = p0 + 16 * w + q;
      if (i >= n) break;
      const uint64_t bb = i / gs::kBlockBases;
      const uint32_t rr = (uint32_t)(i - bb * gs::kBlockBases);
      uint32_t c = 0;
      for (uint32_t s = 1; s < 4; ++s) {
        const uint4 e = occ[bb * 4 + s];
        const uint32_t wd = rr < 32 ? e.y : (rr < 64 ? e.z : e.w);
        if ((wd >> (rr & 31)) & 1u) c = s;
      }
      word |= c << ((15 - q) << 1);
    }
    body[j * 8 + 4 + w] = word;
  }
}

struct Dev {
  std::vector<void*> ptrs;
  template <typename T>
  int alloc(T** p, uint64_t count) {
    *p = nullptr;
    if (hipMalloc((void**)p, count * sizeof(T) + 64) != hipSuccess) return GS_ENOMEM;
    ptrs.push_back(*p);
    return GS_OK;
  }
  void release(void* p) {
    for (auto& q : ptrs)
      if (q == p) {
        (void)hipFree(q);
        q = nullptr;
      }
  }
  ~Dev() {
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  }
};

#define CK(x)                          \
  do {                                 \
    if ((x) != hipSuccess) return GS_EHIP; \
  } while (0)
#define RC(x)               \
  do {                      \
    int rc_ = (x);          \
    if (rc_) return rc_;    \
  } while (0)

struct Temp {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t b) {
    if (b <= cap) return GS_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, b + 256) != hipSuccess) return GS_ENOMEM;
    cap = b + 256;
    return GS_OK;
  }
  ~Temp() {
    if (p) (void)hipFree(p);
  }
};

int build_sa(const uint8_t* dT, uint64_t n, hipStream_t st, Dev& D, uint32_t** sa_out, double* t_sort,
             int* rounds_out) {
  const uint64_t m = n + 1;
  uint64_t *kA, *kB;
  uint32_t *sa, *rank, *vA, *vB, *rows, *rows2;
  uint8_t* flag;
  uint64_t* d_count;
  RC(D.alloc(&kA, m));
  RC(D.alloc(&kB, m));
  RC(D.alloc(&sa, m));
  RC(D.alloc(&rank, m));
  RC(D.alloc(&vA, m));
  RC(D.alloc(&vB, m));
  RC(D.alloc(&rows, m));
  RC(D.alloc(&rows2, m));
  RC(D.alloc(&flag, m));
  RC(D.alloc(&d_count, 1));
  Temp tmp;
  const double t0 = wall();
  hipLaunchKernelGGL(make_keys, dim3(grid_for(m, 256)), dim3(256), 0, st, dT, n, kA, vA);
  {
    rocprim::double_buffer<uint64_t> K(kA, kB);
    rocprim::double_buffer<uint32_t> V(vA, vB);
    size_t bytes = 0;
    CK(rocprim::radix_sort_pairs(nullptr, bytes, K, V, (size_t)m, 0, 3 * kKeyChars, st));
    RC(tmp.ensure(bytes));
    bytes = tmp.cap;
    CK(rocprim::radix_sort_pairs(tmp.p, bytes, K, V, (size_t)m, 0, 3 * kKeyChars, st));
    // group heads over all rows
    uint32_t* hrow = rows2;
    hipLaunchKernelGGL(head_rows, dim3(grid_for(m, 256)), dim3(256), 0, st, K.current(), (const uint32_t*)nullptr,
                       m, hrow);
    uint32_t* grp = (uint32_t*)K.alternate();
    size_t sb = 0;
    CK(rocprim::inclusive_scan(nullptr, sb, hrow, grp, (size_t)m, rocprim::maximum<uint32_t>(), st));
    RC(tmp.ensure(sb));
    sb = tmp.cap;
    CK(rocprim::inclusive_scan(tmp.p, sb, hrow, grp, (size_t)m, rocprim::maximum<uint32_t>(), st));
    hipLaunchKernelGGL(apply_groups, dim3(grid_for(m, 256)), dim3(256), 0, st, (const uint32_t*)nullptr,
                       V.current(), grp, m, sa, rank, flag);
  }
  // rows of unresolved groups: rows[q] = q for flagged q
  {
    rocprim::counting_iterator<uint32_t> it(0);
    size_t sb = 0;
    CK(rocprim::select(nullptr, sb, it, flag, rows, d_count, (size_t)m, st));
    RC(tmp.ensure(sb));
    sb = tmp.cap;
    CK(rocprim::select(tmp.p, sb, it, flag, rows, d_count, (size_t)m, st));
  }
  uint64_t u = 0;
  CK(hipMemcpyAsync(&u, d_count, 8, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  uint64_t H = kKeyChars;
  int rounds = 0;
  while (u > 0) {
    if (H > n) return GS_EINVAL;  // cannot happen for a text with an end marker
    ++rounds;
    hipLaunchKernelGGL(round_keys, dim3(grid_for(u, 256)), dim3(256), 0, st, rows, u, sa, rank, H, n, kA, vA);
    rocprim::double_buffer<uint64_t> K(kA, kB);
    rocprim::double_buffer<uint32_t> V(vA, vB);
    size_t bytes = 0;
    CK(rocprim::radix_sort_pairs(nullptr, bytes, K, V, (size_t)u, 0, 64, st));
    RC(tmp.ensure(bytes));
    bytes = tmp.cap;
    CK(rocprim::radix_sort_pairs(tmp.p, bytes, K, V, (size_t)u, 0, 64, st));
    uint32_t* hrow = rows2;
    hipLaunchKernelGGL(head_rows, dim3(grid_for(u, 256)), dim3(256), 0, st, K.current(), rows, u, hrow);
    uint32_t* grp = (uint32_t*)K.alternate();
    size_t sb = 0;
    CK(rocprim::inclusive_scan(nullptr, sb, hrow, grp, (size_t)u, rocprim::maximum<uint32_t>(), st));
    RC(tmp.ensure(sb));
    sb = tmp.cap;
    CK(rocprim::inclusive_scan(tmp.p, sb, hrow, grp, (size_t)u, rocprim::maximum<uint32_t>(), st));
    hipLaunchKernelGGL(apply_groups, dim3(grid_for(u, 256)), dim3(256), 0, st, rows, V.current(), grp, u, sa, rank,
                       flag);
    sb = 0;
    CK(rocprim::select(nullptr, sb, rows, flag, rows2, d_count, (size_t)u, st));
    RC(tmp.ensure(sb));
    sb = tmp.cap;
    CK(rocprim::select(tmp.p, sb, rows, flag, rows2, d_count, (size_t)u, st));
    CK(hipMemcpyAsync(&u, d_count, 8, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    uint32_t* t = rows;
    rows = rows2;
    rows2 = t;
    H *= 2;
  }
  CK(hipStreamSynchronize(st));
  *t_sort = wall() - t0;
  *rounds_out = rounds;
  // keep only the suffix array
  D.release(kA);
  D.release(kB);
  D.release(rank);
  D.release(vA);
  D.release(vB);
  D.release(rows);
  D.release(rows2);
  D.release(flag);
  *sa_out = sa;
  return GS_OK;
}

int write_files(const char* prefix, const gs_index* X, const uint32_t* d_sa, uint64_t n, hipStream_t st,
                Dev& D) {
  const uint64_t n64 = (n + 63) / 64;
  const uint64_t words = gs_bwt_words(n);
  uint32_t* body;
  RC(D.alloc(&body, n64 * 8 + 8));
  hipLaunchKernelGGL(ref_blocks64, dim3(grid_for(n64, 256)), dim3(256), 0, st, (const uint4*)X->occ,
                     (uint32_t)n, n64, body);
  std::vector<uint32_t> h(n64 * 8 + 8);
  CK(hipMemcpyAsync(h.data(), body, n64 * 8 * 4, hipMemcpyDeviceToHost, st));
  std::vector<uint32_t> hs(X->n_sa);
  CK(hipMemcpyAsync(hs.data(), X->sa, X->n_sa * 4, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  (void)d_sa;
  // the trailing counts block follows the last (possibly partial) block's words
  const uint64_t last_words = n64 ? (n - (n64 - 1) * 64 + 15) / 16 : 0;
  const uint64_t used = n64 ? (n64 - 1) * 8 + 4 + last_words : 0;
  for (uint32_t c = 0; c < 4; ++c) h[used + c] = X->L2[c + 1] - X->L2[c];
  if (used + 4 != words) return GS_EINVAL;
  size_t pl = strlen(prefix);
  std::vector<char> fn(pl + 8);
  memcpy(fn.data(), prefix, pl);
  strcpy(fn.data() + pl, ".bwt");
  FILE* fp = fopen(fn.data(), "wb");
  if (!fp) return GS_EIO;
  bool ok = fwrite(&X->primary, 4, 1, fp) == 1 && fwrite(X->L2 + 1, 4, 4, fp) == 4 &&
            fwrite(h.data(), 4, words, fp) == words;
  ok = (fclose(fp) == 0) && ok;
  if (!ok) return GS_EIO;
  strcpy(fn.data() + pl, ".sa");
  fp = fopen(fn.data(), "wb");
  if (!fp) return GS_EIO;
  const uint32_t n32 = (uint32_t)n;
  ok = fwrite(&X->primary, 4, 1, fp) == 1 && fwrite(X->L2 + 1, 4, 4, fp) == 4 &&
       fwrite(&X->sa_intv, 4, 1, fp) == 1 && fwrite(&n32, 4, 1, fp) == 1 &&
       fwrite(hs.data() + 1, 4, X->n_sa - 1, fp) == X->n_sa - 1;
  ok = (fclose(fp) == 0) && ok;
  return ok ? GS_OK : GS_EIO;
}

}  // namespace

extern "C" GS_API int gs_index_build(const uint8_t* codes, uint64_t n, uint32_t sa_intv, int device,
                                     const char* write_prefix, gs_index** out) {
  if (!codes || !out || n == 0 || n >= 0xFFFFFFFEull || sa_intv == 0) return GS_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return GS_ENODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess || strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return GS_ENODEV;
  if (hipSetDevice(device) != hipSuccess) return GS_ENODEV;
  gs_index* X = new gs_index();
  X->device = device;
  X->seq_len = (uint32_t)n;
  X->sa_intv = sa_intv;
  X->n_sa = (n + sa_intv) / sa_intv;
  X->n_blocks = (n + gs::kBlockBases - 1) / gs::kBlockBases;
  auto fail = [&](int rc) {
    (void)hipFree(X->occ);
    (void)hipFree(X->sa);
    (void)hipFree(X->kmer);
    (void)hipFree(X->kext);
    if (X->stream) (void)hipStreamDestroy(X->stream);
    delete X;
    return rc;
  };
  if (hipStreamCreateWithFlags(&X->stream, hipStreamNonBlocking) != hipSuccess) return fail(GS_EHIP);
  hipStream_t st = X->stream;
  const double t0 = wall();
  int rc = GS_OK;
  {
    Dev D;
    uint8_t* dT;
    if ((rc = D.alloc(&dT, n))) return fail(rc);
    for (uint64_t i = 0; i < n; ++i)
      if (codes[i] > 3) return fail(GS_EINVAL);
    if (hipMemcpyAsync(dT, codes, n, hipMemcpyHostToDevice, st) != hipSuccess) return fail(GS_EHIP);
    uint32_t* d_sa = nullptr;
    double t_sort = 0;
    int rounds = 0;
    if ((rc = build_sa(dT, n, st, D, &d_sa, &t_sort, &rounds))) return fail(rc);
    const double t1 = wall();
    // primary, BWT blocks, counts, L2
    uint32_t* d_prim;
    uint32_t* bcnt;
    uint32_t* bscan;
    const uint64_t nb = X->n_blocks;
    if ((rc = D.alloc(&d_prim, 1)) || (rc = D.alloc(&bcnt, 4 * nb)) || (rc = D.alloc(&bscan, 4 * nb)))
      return fail(rc);
    if (hipMalloc(&X->occ, (nb + 1) * 4 * sizeof(uint4)) != hipSuccess) return fail(GS_ENOMEM);
    if (hipMalloc(&X->sa, X->n_sa * sizeof(uint32_t)) != hipSuccess) return fail(GS_ENOMEM);
    X->device_bytes = (nb + 1) * 4 * sizeof(uint4) + X->n_sa * sizeof(uint32_t);
    hipLaunchKernelGGL(find_primary, dim3(grid_for(n + 1, 256)), dim3(256), 0, st, d_sa, n + 1, d_prim);
    if (hipMemcpyAsync(&X->primary, d_prim, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(GS_EHIP);
    hipLaunchKernelGGL(bwt_blocks96, dim3(grid_for(nb, 256)), dim3(256), 0, st, dT, d_sa, (uint32_t)n, X->primary,
                       nb, X->occ, bcnt);
    if (hipMemsetAsync(X->occ + nb * 4, 0, 4 * sizeof(uint4), st) != hipSuccess) return fail(GS_EHIP);
    {
      Temp tmp;
      for (uint32_t c = 0; c < 4; ++c) {
        size_t sb = 0;
        if (rocprim::exclusive_scan(nullptr, sb, bcnt + c * nb, bscan + c * nb, 0u, (size_t)nb,
                                    rocprim::plus<uint32_t>(), st) != hipSuccess)
          return fail(GS_EHIP);
        if ((rc = tmp.ensure(sb))) return fail(rc);
        sb = tmp.cap;
        if (rocprim::exclusive_scan(tmp.p, sb, bcnt + c * nb, bscan + c * nb, 0u, (size_t)nb,
                                    rocprim::plus<uint32_t>(), st) != hipSuccess)
          return fail(GS_EHIP);
      }
    }
    hipLaunchKernelGGL(set_block_counts, dim3(grid_for(nb, 256)), dim3(256), 0, st, bscan, nb, X->occ);
    hipLaunchKernelGGL(sample_sa, dim3(grid_for(X->n_sa, 256)), dim3(256), 0, st, d_sa, X->n_sa, sa_intv, X->sa);
    std::vector<uint32_t> tail(8);
    if (hipMemcpyAsync(tail.data(), bscan + 0 * nb + nb - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(tail.data() + 1, bscan + 1 * nb + nb - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(tail.data() + 2, bscan + 2 * nb + nb - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(tail.data() + 3, bscan + 3 * nb + nb - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(tail.data() + 4, bcnt + 0 * nb + nb - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(tail.data() + 5, bcnt + 1 * nb + nb - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(tail.data() + 6, bcnt + 2 * nb + nb - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(tail.data() + 7, bcnt + 3 * nb + nb - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(GS_EHIP);
    X->L2[0] = 0;
    for (uint32_t c = 0; c < 4; ++c) X->L2[c + 1] = X->L2[c] + tail[c] + tail[4 + c];
    if (X->L2[4] != n) return fail(GS_EINVAL);
    const double t2 = wall();
    if (write_prefix && (rc = write_files(write_prefix, X, d_sa, n, st, D))) return fail(rc);
    const double t3 = wall();
    X->t_upload = t1 - t0 - t_sort;
    X->t_sa = t_sort;
    X->sa_rounds = rounds;
    X->t_repack = t2 - t1;
    X->t_write = t3 - t2;
  }
  const double t4 = wall();
  if ((rc = gs_build_kmer_table(X))) return fail(rc);
  X->t_table = wall() - t4;
  (void)gs::noalloc_setup(device);  // a failed lookup makes gs_seed_noalloc return GS_EHIP
  *out = X;
  return GS_OK;
}
