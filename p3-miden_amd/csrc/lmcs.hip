// Lifted Merkle tree (LMCS) on gfx950: leaf sponge hashing with lifting,
// 2-to-1 compression layers, row openings and batch-opening hints.
//
// Replaces LiftedMerkleTree::build_with_alignment
// (p3-miden-lmcs/src/lifted_tree.rs:244-310), build_leaf_states_upsampled
// (:385-437), absorb_matrix (:447-483), compress_uniform (:494-539),
// rows (:127-138) and prove_batch (:157-224).
//
// Lifting without materialising upsampled rows: matrices are processed in
// height groups (ascending).  The kernel for a group of height h runs one lane
// per leaf state i < h; it starts from the previous group's state at index
// i >> log2(h / h_prev) — the reference's state duplication, done as an index
// map — absorbs row i of every matrix of that height, and writes the state
// (or, for the last group, the 4-word digest).  Every LDE element is read
// once and no state is duplicated in memory.
//
// Sponge: overwrite mode, rate 8 of width 12 for Poseidon2 (StatefulSponge,
// p3-miden-stateful-hasher/src/field_sponge.rs:164-187), rate 17 of width 25 for
// the u64-serialising Keccak sponge; a partial final chunk is zero-filled.
// Compression: TruncatedPermutation (perm(l || r || 0)[0..4]) for Poseidon2 and
// a single-block PaddingFreeSponge over KeccakF for Keccak (both leave the
// remaining lanes zero).
#include <algorithm>
#include <cstring>
#include <set>

#include "internal.hpp"
#include "p2_fast.hpp"
#include "p2_coop.hpp"

namespace p3g {

constexpr int LEAF_T = 128;  // default lanes per workgroup for leaf hashing

struct LeafMat {
  const uint64_t* ptr;
  uint64_t width;
};

struct LeafArgs {
  const LeafMat* mats;
  uint32_t nmats;
  const uint64_t* prev_states;  // nullptr for the first group
  uint32_t prev_shift;
  uint64_t* out_states;   // non-final groups
  uint64_t* out_digests;  // final group
  uint64_t h;
  const P2Consts* p2;
};

// LDS per wave (LEAF_WAVE_WORDS = 12 x 64 words): words 0..7 of its lanes are
// a prefetch slot in the layout a direct global-to-LDS load writes (16 bytes
// per lane per instruction: piece q of lane l at q * 128 + 2 l), words 8..11
// the permutation's saved capacity (word k of lane l at k * 64 + l).  The
// rare exact redo lays the whole input out as word k of lane l at k * 64 + l
// over the same 768 words, so it needs no more LDS.
constexpr uint32_t LEAF_WAVE_WORDS = 12 * 64;

template <int T>
__device__ __forceinline__ void p2_leaf_body(const LeafArgs& a, uint64_t i, uint64_t* wave_lds, uint32_t lane) {
  if (i >= a.h) return;
  uint64_t st[12];
  if (a.prev_states) {
    const uint64_t* ps = a.prev_states + (i >> a.prev_shift) * 12;
#pragma unroll
    for (int k = 0; k < 12; k++) st[k] = ps[k];
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) st[k] = 0;
  }
  // one chunk of up to 8 words per iteration: overwrite the rate, zero-fill a
  // partial chunk, permute (a single permutation call site keeps the code small)
  uint32_t m = 0;
  uint64_t j = 0;
  // A 128-byte line holds two chunks.  When a chunk starts a line and the next
  // chunk is whole, the next chunk is loaded straight into the wave's LDS slot
  // with the same fetch, so each line leaves the L2 once instead of twice, one
  // permutation apart.
  bool have_next = false;
  const uint4* slot = reinterpret_cast<const uint4*>(wave_lds) + lane;  // piece q at slot[64 q]
  while (m < a.nmats) {
    const uint64_t w = a.mats[m].width;
    if (j >= w) {
      m++;
      j = 0;
      continue;
    }
    const uint64_t* row = a.mats[m].ptr + i * w + j;
    const uint64_t rem = w - j;
    if (have_next) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint4 v = slot[64 * k];
        st[2 * k] = ((uint64_t)v.y << 32) | v.x;
        st[2 * k + 1] = ((uint64_t)v.w << 32) | v.z;
      }
      have_next = false;
    } else if (rem >= 8 && ((((uintptr_t)row) & 15) == 0)) {
      const uint4* r4 = reinterpret_cast<const uint4*>(row);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        uint4 v = r4[k];
        st[2 * k] = ((uint64_t)v.y << 32) | v.x;
        st[2 * k + 1] = ((uint64_t)v.w << 32) | v.z;
      }
      if (rem >= 16 && ((((uintptr_t)row) & 127) == 0)) {
#pragma unroll
        for (int k = 0; k < 4; k++)
          __builtin_amdgcn_global_load_lds((const void*)(r4 + 4 + k),
                                           (__attribute__((address_space(3))) void*)(wave_lds + 128 * k), 16, 0, 0);
        have_next = true;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) st[k] = (uint64_t)k < rem ? row[k] : 0;
    }
#pragma unroll
    for (int k = 8; k < 12; k++) wave_lds[k * 64 + lane] = st[k];
    // the constants pointer made opaque per chunk: the compiler would otherwise
    // hoist the 118 round constants out of the loop into vector registers
    const P2Consts* pc = a.p2;
    asm volatile("" : "+s"(pc));
    // TRY products; the whole wave redoes the permutation exactly from its input
    // (this chunk, re-read from the row, and the saved capacity) when any lane
    // flagged a product borrow (about 2^-32 per product)
    uint64_t flag = 0;
    p2_permute_dev<true>(st, pc, flag);
    if (__builtin_expect(flag != 0, 0)) {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // a prefetch into the slot has landed
      // the chunk's address formed again here, so that it is not held across the permutation
      uint64_t ii = i;
      asm volatile("" : "+v"(ii));
      const uint64_t* row2 = a.mats[m].ptr + ii * w + j;
#pragma unroll
      for (int k = 0; k < 8; k++) wave_lds[k * 64 + lane] = (uint64_t)k < w - j ? row2[k] : 0;
      p2_permute_lds<64>(wave_lds + lane, pc);
#pragma unroll
      for (int k = 0; k < 12; k++) st[k] = wave_lds[k * 64 + lane];
      have_next = false;  // the slot now holds the redo's state
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) st[k] = canon(st[k]);
    }
    j += 8;
  }
  if (a.out_digests) {
#pragma unroll
    for (int k = 0; k < 4; k++) a.out_digests[i * 4 + k] = st[k];
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) a.out_states[i * 12 + k] = st[k];
  }
}

template <int T>
__global__ void __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(5))) p2_leaf_kernel(LeafArgs a) {
  __shared__ uint64_t lds[(T / 64) * LEAF_WAVE_WORDS];
  p2_leaf_body<T>(a, (uint64_t)blockIdx.x * T + threadIdx.x, lds + (threadIdx.x >> 6) * LEAF_WAVE_WORDS,
                  threadIdx.x & 63);
}

// Poseidon2 leaf hashing of a short matrix group: one 16-lane row per leaf
// (the cooperative permutation); same sponge as p2_leaf_kernel.
constexpr uint64_t COOP_LEAF_MAX = 1u << 14;
__device__ __forceinline__ void p2_leaf_coop_body(const LeafArgs& a, uint64_t i, uint32_t l) {
  if (i >= a.h) return;  // uniform over the 16-lane row
  uint64_t x = 0;
  if (a.prev_states && l < 12) x = a.prev_states[(i >> a.prev_shift) * 12 + l];
  for (uint32_t m = 0; m < a.nmats; m++) {
    const uint64_t w = a.mats[m].width;
    const uint64_t* row = a.mats[m].ptr + i * w;
    for (uint64_t j = 0; j < w; j += 8) {
      // overwrite the rate, zero-filling past the row's end
      if (l < 8) x = j + l < w ? row[j + l] : 0;
      x = p2_permute_coop(x, l, a.p2);
    }
  }
  if (a.out_digests) {
    if (l < 4) a.out_digests[i * 4 + l] = x;
  } else if (l < 12) {
    a.out_states[i * 12 + l] = x;
  }
}

__global__ void __launch_bounds__(256) p2_leaf_coop_kernel(LeafArgs a) {
  p2_leaf_coop_body(a, ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 4, threadIdx.x & 15);
}

// One rate block (17 words) of each of a wave's 64 rows.
constexpr uint32_t KECCAK_WAVE_WORDS = 64 * 17;

// Row i of the leaf matrices, hashed by lane `lane` of its wave (rows
// i - lane .. i - lane + 63).  Each Keccak-f absorbs the next 17 words of the
// row (overwrite mode; a partial last block is zero-filled).  The wave loads
// its 64 rows' next blocks together: load t (0..16) gives lane l word
// q = 64 t + l of the wave's block, word q % 17 of row q / 17, so each load
// instruction reads a few contiguous row segments instead of 64 words 8 bytes
// wide from 64 rows one row width apart.  The block goes through the wave's
// LDS slot `wl` (written as q, read back as row l's 17 words) and the next
// block's loads are in flight during the current permutation.  Rows past the
// height read zeros and store nothing.
__device__ __forceinline__ void keccak_leaf_body(const LeafArgs& a, uint64_t i, uint64_t* wl, uint32_t lane) {
  const uint64_t i0 = i - lane;
  const bool live = i < a.h;
  uint64_t st[25];
  if (a.prev_states && live) {
    const uint64_t* ps = a.prev_states + (i >> a.prev_shift) * 25;
#pragma unroll
    for (int k = 0; k < 25; k++) st[k] = ps[k];
  } else {
#pragma unroll
    for (int k = 0; k < 25; k++) st[k] = 0;
  }
  for (uint32_t m = 0; m < a.nmats; m++) {
    const uint64_t w = a.mats[m].width;
    const uint64_t* base = a.mats[m].ptr + i0 * w;
    uint64_t v[17];
    const bool all_rows = i0 + 64 <= a.h;  // wave-uniform
    auto load = [&](uint64_t j) {
      if (all_rows && j + 17 <= w) {
#pragma unroll
        for (int t = 0; t < 17; t++) {
          const uint32_t q = 64u * t + lane, r = q / 17u, k = q - 17u * r;
          v[t] = base[(uint64_t)r * w + j + k];
        }
      } else {
#pragma unroll
        for (int t = 0; t < 17; t++) {
          const uint32_t q = 64u * t + lane, r = q / 17u, k = q - 17u * r;
          v[t] = (i0 + r < a.h && j + k < w) ? base[(uint64_t)r * w + j + k] : 0;
        }
      }
    };
    if (w) load(0);
    for (uint64_t j = 0; j < w; j += 17) {
#pragma unroll
      for (int t = 0; t < 17; t++) wl[64 * t + lane] = v[t];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < 17; k++) st[k] = wl[17 * lane + k];
      __builtin_amdgcn_wave_barrier();
      if (j + 17 < w) load(j + 17);
      keccak_f(st);
    }
  }
  if (!live) return;
  if (a.out_digests) {
#pragma unroll
    for (int k = 0; k < 4; k++) a.out_digests[i * 4 + k] = st[k];
  } else {
#pragma unroll
    for (int k = 0; k < 25; k++) a.out_states[i * 25 + k] = st[k];
  }
}

template <int T>
__global__ void __launch_bounds__(T) keccak_leaf_kernel(LeafArgs a) {
  __shared__ uint64_t lds[(T / 64) * KECCAK_WAVE_WORDS];
  keccak_leaf_body(a, (uint64_t)blockIdx.x * T + threadIdx.x, lds + (threadIdx.x >> 6) * KECCAK_WAVE_WORDS,
                   threadIdx.x & 63);
}

// ---- job-indexed forms for batches of same-shaped trees (blockIdx.y = job).
// The Poseidon2 constants come as a kernel argument, not through the job's
// descriptor: a pointer read from memory kept them in vector registers
// (p2_leaf_many_kernel: 112 VGPRs, 4 waves per SIMD; 82 and 5 this way).
__global__ void __launch_bounds__(128) p2_leaf_many_kernel(const LeafArgs* __restrict__ args,
                                                        const P2Consts* __restrict__ p2) {
  __shared__ uint64_t lds[2 * LEAF_WAVE_WORDS];
  LeafArgs a = args[blockIdx.y];
  a.p2 = p2;
  p2_leaf_body<128>(a, (uint64_t)blockIdx.x * 128 + threadIdx.x, lds + (threadIdx.x >> 6) * LEAF_WAVE_WORDS,
                    threadIdx.x & 63);
}
__global__ void __launch_bounds__(256) p2_leaf_coop_many_kernel(const LeafArgs* __restrict__ args,
                                                             const P2Consts* __restrict__ p2) {
  LeafArgs a = args[blockIdx.y];
  a.p2 = p2;
  p2_leaf_coop_body(a, ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 4, threadIdx.x & 15);
}
__global__ void __launch_bounds__(128) keccak_leaf_many_kernel(const LeafArgs* __restrict__ args) {
  __shared__ uint64_t lds[2 * KECCAK_WAVE_WORDS];
  keccak_leaf_body(args[blockIdx.y], (uint64_t)blockIdx.x * 128 + threadIdx.x,
                   lds + (threadIdx.x >> 6) * KECCAK_WAVE_WORDS, threadIdx.x & 63);
}

constexpr int CMP_T = 128;

__device__ __forceinline__ void compress_one(uint32_t hash, const uint64_t* l, const uint64_t* r, uint64_t* out,
                                             const P2Consts* p2) {
  if (hash == HASH_P2) {
    uint64_t s[12];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      s[k] = l[k];
      s[4 + k] = r[k];
      s[8 + k] = 0;
    }
    p2_permute_fast(s, p2);
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = s[k];
  } else {
    uint64_t s[25];
#pragma unroll
    for (int k = 0; k < 25; k++) s[k] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      s[k] = l[k];
      s[4 + k] = r[k];
    }
    keccak_f(s);
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = s[k];
  }
}

// save: this thread's slots (stride CMP_T) of a block-wide LDS array (Poseidon2)
template <uint32_t HASH>
__device__ __forceinline__ void compress_layer_body(const uint64_t* __restrict__ prev, uint64_t* __restrict__ next,
                                                    uint64_t len, const P2Consts* p2, uint64_t i, uint64_t* save) {
  if (i >= len) return;
  uint64_t l[4], r[4], o[4];
  const uint64_t* src = prev + i * 8;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    l[k] = src[k];
    r[k] = src[4 + k];
  }
  if constexpr (HASH == HASH_P2) {
    uint64_t st[12];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      st[k] = l[k];
      st[4 + k] = r[k];
      st[8 + k] = 0;
    }
    p2_permute_saved<CMP_T>(st, p2, save);
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = st[k];
  } else {
    compress_one(HASH, l, r, o, p2);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) next[i * 4 + k] = o[k];
}

template <uint32_t HASH>
__global__ void __launch_bounds__(CMP_T) compress_layer_kernel(const uint64_t* __restrict__ prev,
                                                               uint64_t* __restrict__ next, uint64_t len,
                                                               const P2Consts* p2) {
  __shared__ uint64_t save[HASH == HASH_P2 ? 12 * CMP_T : 1];
  compress_layer_body<HASH>(prev, next, len, p2, (uint64_t)blockIdx.x * CMP_T + threadIdx.x, save + threadIdx.x);
}

// one layer of a batch of same-shaped trees: job j reads layer[j].prev, writes layer[j].next
struct LayerJob {
  const uint64_t* prev;
  uint64_t* next;
};
template <uint32_t HASH>
__global__ void __launch_bounds__(CMP_T) compress_layer_many_kernel(const LayerJob* __restrict__ jobs, uint64_t len,
                                                                    const P2Consts* p2) {
  __shared__ uint64_t save[HASH == HASH_P2 ? 12 * CMP_T : 1];
  const LayerJob j = jobs[blockIdx.y];
  compress_layer_body<HASH>(j.prev, j.next, len, p2, (uint64_t)blockIdx.x * CMP_T + threadIdx.x, save + threadIdx.x);
}

// Poseidon2 layer with few nodes: one 16-lane row per node (the cooperative
// permutation), for levels too small to fill the device one node per lane.
constexpr uint64_t COOP_LAYER_MAX = 1u << 14;
__device__ __forceinline__ void compress_layer_coop_body(const uint64_t* __restrict__ prev, uint64_t* __restrict__ next,
                                                         uint64_t len, const P2Consts* p2, uint64_t node, uint32_t l) {
  if (node >= len) return;  // uniform over the 16-lane row
  uint64_t x = l < 8 ? prev[node * 8 + l] : 0;
  x = p2_permute_coop(x, l, p2);
  if (l < 4) next[node * 4 + l] = x;
}
__global__ void __launch_bounds__(256) compress_layer_coop_kernel(const uint64_t* __restrict__ prev,
                                                                  uint64_t* __restrict__ next, uint64_t len,
                                                                  const P2Consts* p2) {
  compress_layer_coop_body(prev, next, len, p2, ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 4, threadIdx.x & 15);
}
__global__ void __launch_bounds__(256) compress_layer_coop_many_kernel(const LayerJob* __restrict__ jobs, uint64_t len,
                                                                       const P2Consts* p2) {
  const LayerJob j = jobs[blockIdx.y];
  compress_layer_coop_body(j.prev, j.next, len, p2, ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 4,
                           threadIdx.x & 15);
}

// Remaining levels once a layer has at most 2*TOP_T digests: one workgroup,
// levels staged through LDS, every level written out.
constexpr int TOP_T = 256;
template <uint32_t HASH>
__device__ __forceinline__ void compress_top_body(uint64_t* digests, uint64_t off, uint64_t len, const P2Consts* p2) {
  __shared__ uint64_t buf[2][TOP_T * 2 * 4];
  const uint32_t tid = threadIdx.x;
  for (uint64_t e = tid; e < len * 4; e += TOP_T) buf[0][e] = digests[off * 4 + e];
  __syncthreads();
  int cur = 0;
  while (len > 1) {
    uint64_t nl = len / 2;
    uint64_t noff = off + len;
    if (tid < nl) {
      uint64_t o[4];
      compress_one(HASH, &buf[cur][tid * 8], &buf[cur][tid * 8 + 4], o, p2);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        buf[cur ^ 1][tid * 4 + k] = o[k];
        digests[(noff + tid) * 4 + k] = o[k];
      }
    }
    __syncthreads();
    cur ^= 1;
    off = noff;
    len = nl;
  }
}

template <uint32_t HASH>
__global__ void __launch_bounds__(TOP_T) compress_top_kernel(uint64_t* digests, uint64_t off, uint64_t len,
                                                             const P2Consts* p2) {
  compress_top_body<HASH>(digests, off, len, p2);
}
// batch of same-shaped trees: one workgroup per tree (digest arrays per job)
template <uint32_t HASH>
__global__ void __launch_bounds__(TOP_T) compress_top_many_kernel(uint64_t* const* __restrict__ digests, uint64_t off,
                                                                  uint64_t len, const P2Consts* p2) {
  compress_top_body<HASH>(digests[blockIdx.x], off, len, p2);
}

// Poseidon2 variant of the tree top: every level with the 16-lane cooperative
// permutation (whose dependent chain is about a fifth of the one-lane form's),
// one node per 16-lane row.
constexpr int TOPC_T = 1024;
// The top starts once a layer holds at most this many digests, so that every
// level in it is a single pass of the 64 rows.  Measured: a level inside the
// top costs about as much as a separate cooperative layer launch (~17 us),
// while a wider level inside it (two passes, or one node per lane) costs
// ~50 us.
constexpr uint64_t TOPC_MAX_LEN = 2 * (TOPC_T / 16);
__host__ __device__ constexpr uint64_t top_max_len(uint32_t hash) {
  return hash == HASH_P2 ? TOPC_MAX_LEN : 2 * TOP_T;
}
// Levels above a layer of `len` digests at layer offset `off`, in slices of
// `slen` digests (a power of two, at most TOPC_MAX_LEN): workgroup b compresses
// digests b*slen .. b*slen + slen - 1 up log2(slen) levels, staged through LDS,
// and writes every level where the tree stores it (layers lie one after
// another, so the layer above starts at off + len).  One slice (slen == len)
// is the tree's top; many slices take the levels of a wider layer in one
// launch instead of one launch per level, each of which costs a kernel gap.
__device__ __forceinline__ void compress_slice_p2_body(uint64_t* digests, uint64_t off, uint64_t len, uint64_t slen,
                                                       uint64_t b, const P2Consts* p2) {
  __shared__ uint64_t buf[2][TOPC_MAX_LEN * 4];
  const uint32_t tid = threadIdx.x;
  const uint32_t l = tid & 15, node = tid >> 4;
  for (uint64_t e = tid; e < slen * 4; e += TOPC_T) buf[0][e] = digests[(off + b * slen) * 4 + e];
  __syncthreads();
  int cur = 0;
  while (slen > 1) {
    const uint64_t nl = slen / 2;
    const uint64_t noff = off + len;
    if (node < nl) {  // uniform over the 16-lane row
      uint64_t x = l < 8 ? buf[cur][node * 8 + l] : 0;
      x = p2_permute_coop(x, l, p2);
      if (l < 4) {
        buf[cur ^ 1][node * 4 + l] = x;
        digests[(noff + b * nl + node) * 4 + l] = x;
      }
    }
    __syncthreads();
    cur ^= 1;
    off = noff;
    len /= 2;
    slen = nl;
  }
}
// len <= TOPC_MAX_LEN digests at layer offset off
__device__ __forceinline__ void compress_top_p2_body(uint64_t* digests, uint64_t off, uint64_t len,
                                                     const P2Consts* p2) {
  compress_slice_p2_body(digests, off, len, len, 0, p2);
}

__global__ void __launch_bounds__(TOPC_T) compress_slices_p2_kernel(uint64_t* digests, uint64_t off, uint64_t len,
                                                                    uint64_t slen, const P2Consts* p2) {
  compress_slice_p2_body(digests, off, len, slen, blockIdx.x, p2);
}
__global__ void __launch_bounds__(TOPC_T) compress_top_p2_kernel(uint64_t* digests, uint64_t off, uint64_t len,
                                                                 const P2Consts* p2) {
  compress_top_p2_body(digests, off, len, p2);
}
__global__ void __launch_bounds__(TOPC_T) compress_top_p2_many_kernel(uint64_t* const* __restrict__ digests,
                                                                      uint64_t off, uint64_t len, const P2Consts* p2) {
  compress_top_p2_body(digests[blockIdx.x], off, len, p2);
}

// 16 lanes per piece (most pieces are a digest or a short row): piece
// b = global lane / 16 copies lens[b] words from srcs[b] to out[offs[b]..].
__global__ void __launch_bounds__(256) gather_kernel(const uint64_t* const* __restrict__ srcs,
                                                     const uint64_t* __restrict__ lens,
                                                     const uint64_t* __restrict__ offs, uint64_t np,
                                                     uint64_t* __restrict__ out) {
  const uint64_t b = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 4;
  if (b >= np) return;
  const uint64_t* src = srcs[b];
  const uint64_t n = lens[b], o = offs[b];
  for (uint64_t e = threadIdx.x & 15; e < n; e += 16) out[o + e] = src[e];
}

// ------------------------------------------------------------------ host side

static void launch_leaf(Ctx& ctx, uint32_t hash, const std::vector<DevMat>& group, const uint64_t* prev,
                        uint32_t prev_shift, uint64_t* out_states, uint64_t* out_digests, uint64_t h,
                        hipStream_t s) {
  std::vector<LeafMat> hm(group.size());
  for (size_t i = 0; i < group.size(); i++) hm[i] = LeafMat{group[i].ptr, group[i].width};
  LeafMat* dm = (LeafMat*)ctx.upload(hm.data(), hm.size() * sizeof(LeafMat), s);
  LeafArgs a{};
  a.mats = dm;
  a.nmats = (uint32_t)hm.size();
  a.prev_states = prev;
  a.prev_shift = prev_shift;
  a.out_states = out_states;
  a.out_digests = out_digests;
  a.h = h;
  a.p2 = ctx.dev_p2;
  uint32_t grid = (uint32_t)((h + LEAF_T - 1) / LEAF_T);
  {
  ProfScope ps(ctx, hash == HASH_P2 ? "p2_leaf" : "keccak_leaf", s);
  if (hash == HASH_P2 && h <= COOP_LEAF_MAX)
    hipLaunchKernelGGL(p2_leaf_coop_kernel, dim3((uint32_t)((h * 16 + 255) / 256)), dim3(256), 0, s, a);
  else if (hash == HASH_P2) hipLaunchKernelGGL(p2_leaf_kernel<LEAF_T>, dim3(grid), dim3(LEAF_T), 0, s, a);
  else hipLaunchKernelGGL(keccak_leaf_kernel<LEAF_T>, dim3(grid), dim3(LEAF_T), 0, s, a);
  P3G_CHECK_HIP(hipGetLastError());
  }
  ctx.release(dm);  // stream-ordered: a later owner's work queues behind this launch
}

// All compression layers of a tree whose leaf digests are in place.  Only
// kernel launches (safe inside stream capture); ctx may be null (no timing).
static void compress_layers(Ctx* ctx, uint32_t hash, uint64_t* digests, const std::vector<uint64_t>& layer_off,
                            uint64_t height, const P2Consts* p2, hipStream_t s) {
  uint64_t len = height;
  unsigned l = 0;
  while (len > top_max_len(hash)) {
    uint64_t nl = len / 2;
    if (hash == HASH_P2 && nl <= COOP_LAYER_MAX) {
      // the cooperative levels in slices: down to TOPC_MAX_LEN digests in at
      // most two launches (a layer of <= 2^14 nodes has <= 2^15 digests)
      const uint64_t slen = std::min<uint64_t>(TOPC_MAX_LEN, std::max<uint64_t>(2, len / TOPC_MAX_LEN));
      auto slices = [&] {
        hipLaunchKernelGGL(compress_slices_p2_kernel, dim3((uint32_t)(len / slen)), dim3(TOPC_T), 0, s, digests,
                           layer_off[l], len, slen, p2);
      };
      if (ctx) {
        ProfScope ps(*ctx, "compress_layer", s);
        slices();
      } else {
        slices();
      }
      P3G_CHECK_HIP(hipGetLastError());
      for (uint64_t t = slen; t > 1; t /= 2) {
        len /= 2;
        l++;
      }
      continue;
    }
    uint32_t grid = (uint32_t)((nl + CMP_T - 1) / CMP_T);
    const uint64_t* prev = digests + layer_off[l] * 4;
    uint64_t* next = digests + layer_off[l + 1] * 4;
    auto layer = [&] {
      if (hash == HASH_P2 && nl <= COOP_LAYER_MAX)
        hipLaunchKernelGGL(compress_layer_coop_kernel, dim3((uint32_t)((nl * 16 + 255) / 256)), dim3(256), 0, s, prev,
                           next, nl, p2);
      else if (hash == HASH_P2)
        hipLaunchKernelGGL(compress_layer_kernel<HASH_P2>, dim3(grid), dim3(CMP_T), 0, s, prev, next, nl, p2);
      else
        hipLaunchKernelGGL(compress_layer_kernel<HASH_KECCAK>, dim3(grid), dim3(CMP_T), 0, s, prev, next, nl, p2);
    };
    if (ctx) {
      ProfScope ps(*ctx, "compress_layer", s);
      layer();
    } else {
      layer();
    }
    P3G_CHECK_HIP(hipGetLastError());
    len = nl;
    l++;
  }
  if (len > 1) {
    if (len > top_max_len(hash)) throw ApiError(E_RANGE, "tree top wider than its kernel");
    auto top = [&] {
      if (hash == HASH_P2)
        hipLaunchKernelGGL(compress_top_p2_kernel, dim3(1), dim3(TOPC_T), 0, s, digests, layer_off[l], len, p2);
      else
        hipLaunchKernelGGL(compress_top_kernel<HASH_KECCAK>, dim3(1), dim3(TOP_T), 0, s, digests, layer_off[l], len, p2);
    };
    if (ctx) {
      ProfScope ps(*ctx, "compress_top", s);
      top();
    } else {
      top();
    }
    P3G_CHECK_HIP(hipGetLastError());
  }
}

static void build_layers(Ctx& ctx, Tree& t, hipStream_t s) {
  compress_layers(&ctx, t.hash, t.digests, t.layer_off, t.height, ctx.dev_p2, s);
}

void launch_compress_layers(uint32_t hash, uint64_t* digests, const std::vector<uint64_t>& layer_off,
                            uint64_t height, const P2Consts* p2, hipStream_t s) {
  compress_layers(nullptr, hash, digests, layer_off, height, p2, s);
}

size_t leaf_mat_size() { return sizeof(LeafMat); }
void make_leaf_mat(void* out, const uint64_t* ptr, uint64_t width) {
  LeafMat m{ptr, width};
  memcpy(out, &m, sizeof(m));
}

// Leaf hashing of one height group with device-resident matrix descriptors
// (no host interaction: safe inside stream capture).
void launch_leaf_raw(uint32_t hash, int threads, const void* mats, uint32_t nmats, const uint64_t* prev,
                     uint32_t prev_shift, uint64_t* out_states, uint64_t* out_digests, uint64_t h,
                     const P2Consts* p2, hipStream_t s) {
  LeafArgs a{};
  a.mats = reinterpret_cast<const LeafMat*>(mats);
  a.nmats = nmats;
  a.prev_states = prev;
  a.prev_shift = prev_shift;
  a.out_states = out_states;
  a.out_digests = out_digests;
  a.h = h;
  a.p2 = p2;
  const uint32_t grid = (uint32_t)((h + threads - 1) / threads);
  if (hash == HASH_P2 && h <= COOP_LEAF_MAX) {
    hipLaunchKernelGGL(p2_leaf_coop_kernel, dim3((uint32_t)((h * 16 + 255) / 256)), dim3(256), 0, s, a);
  } else if (hash == HASH_P2) {
    if (threads == 64) hipLaunchKernelGGL(p2_leaf_kernel<64>, dim3(grid), dim3(64), 0, s, a);
    else if (threads == 256) hipLaunchKernelGGL(p2_leaf_kernel<256>, dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(p2_leaf_kernel<128>, dim3(grid), dim3(128), 0, s, a);
  } else {
    if (threads == 64) hipLaunchKernelGGL(keccak_leaf_kernel<64>, dim3(grid), dim3(64), 0, s, a);
    else if (threads == 256) hipLaunchKernelGGL(keccak_leaf_kernel<256>, dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(keccak_leaf_kernel<128>, dim3(grid), dim3(128), 0, s, a);
  }
  P3G_CHECK_HIP(hipGetLastError());
}

Tree* lmcs_build(Ctx& ctx, uint32_t hash, const std::vector<DevMat>& mats, uint32_t alignment, hipStream_t s,
                 bool read_root, const DevMat* salt) {
  if (mats.empty()) throw ApiError(E_EMPTY, "cannot commit empty batch");
  uint64_t prevh = 0;
  for (auto& m : mats) {
    if (m.height == 0 || (m.height & (m.height - 1))) throw ApiError(E_NOT_POW2, "matrix height must be a power of two");
    if (m.height < prevh) throw ApiError(E_UNSORTED, "matrices must be sorted by height");
    prevh = m.height;
  }
  if (salt && (salt->height != prevh || salt->width == 0 || !salt->ptr))
    throw ApiError(E_INVAL, "salt must have the tree height and a non-zero width");
  Tree* t = new Tree();
  if (salt) t->salt = *salt;
  t->ctx = &ctx;
  t->hash = hash;
  t->alignment = alignment < 1 ? 1 : alignment;
  t->mats = mats;
  t->height = mats.back().height;
  t->depth = log2_exact(t->height);
  t->layer_off.resize(t->depth + 1);
  uint64_t off = 0;
  for (unsigned l = 0; l <= t->depth; l++) {
    t->layer_off[l] = off;
    off += t->height >> l;
  }
  t->digests = (uint64_t*)ctx.alloc(off * 4 * sizeof(uint64_t));

  const uint32_t W = hash_width(hash);
  uint64_t* states[2] = {nullptr, nullptr};
  int cur = -1;
  uint64_t cur_h = 0;
  size_t i = 0;
  while (i < mats.size()) {
    size_t j = i;
    while (j < mats.size() && mats[j].height == mats[i].height) j++;
    std::vector<DevMat> group(mats.begin() + i, mats.begin() + j);
    uint64_t h = mats[i].height;
    bool final = (j == mats.size());
    const uint64_t* prev = cur >= 0 ? states[cur] : nullptr;
    uint32_t shift = cur >= 0 ? log2_exact(h / cur_h) : 0;
    int nxt = cur >= 0 ? cur ^ 1 : 0;
    uint64_t* out_states = nullptr;
    if (!final) {
      if (!states[nxt]) states[nxt] = (uint64_t*)ctx.alloc(t->height * W * sizeof(uint64_t));
      out_states = states[nxt];
    }
    // a hiding tree absorbs each leaf's salt after its rows (build_with_alignment:
    // absorb_matrix of the salt matrix, lifted_tree.rs:271-279)
    if (final && salt) group.push_back(*salt);
    launch_leaf(ctx, hash, group, prev, shift, out_states, final ? t->digests : nullptr, h, s);
    cur = nxt;
    cur_h = h;
    i = j;
  }
  build_layers(ctx, *t, s);
  // without the read-back nothing here waits for the stream (the releases
  // below are stream-ordered); t->root is then filled in by the caller
  if (read_root) ctx.read_back(t->root, t->digests + t->layer_off[t->depth] * 4, 32, s);
  ctx.release(states[0]);
  ctx.release(states[1]);
  return t;
}

std::vector<Tree*> lmcs_build_many(Ctx& ctx, uint32_t hash, const std::vector<std::vector<DevMat>>& mats,
                                   uint32_t alignment, hipStream_t s, bool read_roots) {
  const size_t k = mats.size();
  std::vector<Tree*> trees;
  if (k == 0) return trees;
  const std::vector<DevMat>& m0 = mats[0];
  if (m0.empty()) throw ApiError(E_EMPTY, "cannot commit empty batch");
  uint64_t prevh = 0;
  for (auto& m : m0) {
    if (m.height == 0 || (m.height & (m.height - 1))) throw ApiError(E_NOT_POW2, "matrix height must be a power of two");
    if (m.height < prevh) throw ApiError(E_UNSORTED, "matrices must be sorted by height");
    prevh = m.height;
  }
  for (auto& mj : mats) {
    if (mj.size() != m0.size()) throw ApiError(E_INVAL, "batched trees need one shape");
    for (size_t i = 0; i < mj.size(); i++)
      if (mj[i].height != m0[i].height || mj[i].width != m0[i].width) throw ApiError(E_INVAL, "batched trees need one shape");
  }
  const uint64_t H = m0.back().height;
  const unsigned depth = log2_exact(H);
  const uint32_t W = hash_width(hash);
  std::vector<uint64_t*> scratch;
  uint8_t* blob = nullptr;
  std::vector<uint8_t> host;  // host image of the launch descriptors (outlives its upload)
  try {
    std::vector<uint64_t> layer_off(depth + 1);
    uint64_t off = 0;
    for (unsigned l = 0; l <= depth; l++) {
      layer_off[l] = off;
      off += H >> l;
    }
    for (size_t j = 0; j < k; j++) {
      Tree* t = new Tree();
      t->ctx = &ctx;
      t->hash = hash;
      t->alignment = alignment < 1 ? 1 : alignment;
      t->mats = mats[j];
      t->height = H;
      t->depth = depth;
      t->layer_off = layer_off;
      trees.push_back(t);
      t->digests = (uint64_t*)ctx.alloc(off * 4 * sizeof(uint64_t));
    }
    // height groups
    std::vector<std::pair<size_t, size_t>> groups;
    for (size_t i = 0; i < m0.size();) {
      size_t e = i;
      while (e < m0.size() && m0[e].height == m0[i].height) e++;
      groups.emplace_back(i, e);
      i = e;
    }
    std::vector<uint64_t*> states[2];
    if (groups.size() > 1)
      for (int b = 0; b < 2; b++)
        for (size_t j = 0; j < k; j++) {
          uint64_t* p = (uint64_t*)ctx.alloc(H * W * sizeof(uint64_t));
          scratch.push_back(p);
          states[b].push_back(p);
        }
    // descriptor layout: [LeafMat arrays | LeafArgs per group | LayerJob per level | digest ptrs]
    auto reserve = [&](size_t bytes) {
      size_t o = (host.size() + 255) & ~size_t(255);
      host.resize(o + bytes);
      return o;
    };
    std::vector<size_t> args_off(groups.size());
    for (size_t g = 0; g < groups.size(); g++) {
      const size_t nm = groups[g].second - groups[g].first;
      std::vector<size_t> mats_off(k);
      for (size_t j = 0; j < k; j++) mats_off[j] = reserve(nm * sizeof(LeafMat));
      args_off[g] = reserve(k * sizeof(LeafArgs));
      const uint64_t h = m0[groups[g].first].height;
      for (size_t j = 0; j < k; j++) {
        for (size_t q = 0; q < nm; q++) {
          LeafMat lm{mats[j][groups[g].first + q].ptr, mats[j][groups[g].first + q].width};
          memcpy(host.data() + mats_off[j] + q * sizeof(LeafMat), &lm, sizeof(lm));
        }
        LeafArgs a{};
        a.mats = reinterpret_cast<const LeafMat*>(mats_off[j]);  // relocated after upload
        a.nmats = (uint32_t)nm;
        a.prev_states = g ? states[(g - 1) & 1][j] : nullptr;
        a.prev_shift = g ? log2_exact(h / m0[groups[g - 1].first].height) : 0;
        a.out_states = g + 1 < groups.size() ? states[g & 1][j] : nullptr;
        a.out_digests = g + 1 < groups.size() ? nullptr : trees[j]->digests;
        a.h = h;
        a.p2 = ctx.dev_p2;
        memcpy(host.data() + args_off[g] + j * sizeof(LeafArgs), &a, sizeof(a));
      }
    }
    std::vector<size_t> level_off;
    uint64_t len = H;
    unsigned lvl = 0;
    while (len > top_max_len(hash)) {
      const size_t o = reserve(k * sizeof(LayerJob));
      for (size_t j = 0; j < k; j++) {
        LayerJob lj{trees[j]->digests + layer_off[lvl] * 4, trees[j]->digests + layer_off[lvl + 1] * 4};
        memcpy(host.data() + o + j * sizeof(LayerJob), &lj, sizeof(lj));
      }
      level_off.push_back(o);
      len /= 2;
      lvl++;
    }
    const size_t dptr_off = reserve(k * sizeof(uint64_t*));
    for (size_t j = 0; j < k; j++) memcpy(host.data() + dptr_off + j * sizeof(uint64_t*), &trees[j]->digests, 8);
    blob = (uint8_t*)ctx.alloc(host.size());
    // relocate the LeafMat array pointers to device addresses
    for (size_t g = 0; g < groups.size(); g++)
      for (size_t j = 0; j < k; j++) {
        LeafArgs* a = reinterpret_cast<LeafArgs*>(host.data() + args_off[g] + j * sizeof(LeafArgs));
        a->mats = reinterpret_cast<const LeafMat*>(blob + reinterpret_cast<uintptr_t>(a->mats));
      }
    ctx.upload_async(blob, host.data(), host.size(), s);
    {
      ProfScope ps(ctx, hash == HASH_P2 ? "p2_leaf" : "keccak_leaf", s);
      for (size_t g = 0; g < groups.size(); g++) {
        const uint64_t h = m0[groups[g].first].height;
        const LeafArgs* da = reinterpret_cast<const LeafArgs*>(blob + args_off[g]);
        // the cooperative form costs more lanes per permutation: worth it only
        // while the whole launch (all jobs) is too small to fill the device
        if (hash == HASH_P2 && h * k <= COOP_LEAF_MAX)
          hipLaunchKernelGGL(p2_leaf_coop_many_kernel, dim3((uint32_t)((h * 16 + 255) / 256), (uint32_t)k), dim3(256),
                             0, s, da, ctx.dev_p2);
        else if (hash == HASH_P2)
          hipLaunchKernelGGL(p2_leaf_many_kernel, dim3((uint32_t)((h + 127) / 128), (uint32_t)k), dim3(128), 0, s, da,
                             ctx.dev_p2);
        else
          hipLaunchKernelGGL(keccak_leaf_many_kernel, dim3((uint32_t)((h + 127) / 128), (uint32_t)k), dim3(128), 0, s,
                             da);
        P3G_CHECK_HIP(hipGetLastError());
      }
    }
    len = H;
    for (size_t li = 0; li < level_off.size(); li++) {
      const uint64_t nl = len / 2;
      const LayerJob* dj = reinterpret_cast<const LayerJob*>(blob + level_off[li]);
      ProfScope ps(ctx, "compress_layer", s);
      if (hash == HASH_P2 && nl * k <= COOP_LAYER_MAX)
        hipLaunchKernelGGL(compress_layer_coop_many_kernel, dim3((uint32_t)((nl * 16 + 255) / 256), (uint32_t)k),
                           dim3(256), 0, s, dj, nl, ctx.dev_p2);
      else if (hash == HASH_P2)
        hipLaunchKernelGGL(compress_layer_many_kernel<HASH_P2>, dim3((uint32_t)((nl + CMP_T - 1) / CMP_T), (uint32_t)k),
                           dim3(CMP_T), 0, s, dj, nl, ctx.dev_p2);
      else
        hipLaunchKernelGGL(compress_layer_many_kernel<HASH_KECCAK>,
                           dim3((uint32_t)((nl + CMP_T - 1) / CMP_T), (uint32_t)k), dim3(CMP_T), 0, s, dj, nl,
                           ctx.dev_p2);
      P3G_CHECK_HIP(hipGetLastError());
      len = nl;
    }
    if (len > 1) {
      if (len > top_max_len(hash)) throw ApiError(E_RANGE, "tree top wider than its kernel");
      uint64_t* const* dd = reinterpret_cast<uint64_t* const*>(blob + dptr_off);
      ProfScope ps(ctx, "compress_top", s);
      if (hash == HASH_P2)
        hipLaunchKernelGGL(compress_top_p2_many_kernel, dim3((uint32_t)k), dim3(TOPC_T), 0, s, dd,
                           layer_off[level_off.size()], len, ctx.dev_p2);
      else
        hipLaunchKernelGGL(compress_top_many_kernel<HASH_KECCAK>, dim3((uint32_t)k), dim3(TOP_T), 0, s, dd,
                           layer_off[level_off.size()], len, ctx.dev_p2);
      P3G_CHECK_HIP(hipGetLastError());
    }
    // roots (without the read-back the roots stay on the device and nothing
    // here waits for the stream)
    if (read_roots) {
      std::vector<std::pair<const uint64_t*, uint64_t>> pieces;
      for (size_t j = 0; j < k; j++) pieces.emplace_back(trees[j]->digests + layer_off[depth] * 4, 4);
      std::vector<uint64_t> raw;
      gather_to_host(ctx, pieces, raw, s);  // synchronises the stream
      for (size_t j = 0; j < k; j++) memcpy(trees[j]->root, raw.data() + 4 * j, 32);
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);
    for (auto* p : scratch) ctx.release(p);
    ctx.release(blob);
    for (auto* t : trees) delete t;
    throw;
  }
  for (auto* p : scratch) ctx.release(p);
  ctx.release(blob);
  return trees;
}

const uint64_t* gather_pinned(Ctx& ctx, const std::vector<std::pair<const uint64_t*, uint64_t>>& pieces,
                              hipStream_t s) {
  const size_t np = pieces.size();
  uint64_t total = 0;
  for (auto& pc : pieces) total += pc.second;
  // pinned image: [srcs | lens | offs | gathered words]
  uint64_t* hs = ctx.host_stage(3 * np + total + 1);
  if (np == 0 || total == 0) return hs + 3 * np;
  uint64_t acc = 0;
  for (size_t i = 0; i < np; i++) {
    hs[i] = reinterpret_cast<uint64_t>(pieces[i].first);
    hs[np + i] = pieces[i].second;
    hs[2 * np + i] = acc;
    acc += pieces[i].second;
  }
  uint64_t* stage = (uint64_t*)ctx.alloc((3 * np + total) * sizeof(uint64_t));
  try {
    P3G_CHECK_HIP(hipMemcpyAsync(stage, hs, 3 * np * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gather_kernel, dim3((uint32_t)((np * 16 + 255) / 256)), dim3(256), 0, s,
                       (const uint64_t* const*)stage, stage + np, stage + 2 * np, (uint64_t)np, stage + 3 * np);
    P3G_CHECK_HIP(hipGetLastError());
    P3G_CHECK_HIP(hipMemcpyAsync(hs + 3 * np, stage + 3 * np, total * 8, hipMemcpyDeviceToHost, s));
    P3G_CHECK_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    ctx.release(stage);
    throw;
  }
  ctx.release(stage);
  return hs + 3 * np;
}

void gather_to_host(Ctx& ctx, const std::vector<std::pair<const uint64_t*, uint64_t>>& pieces,
                    std::vector<uint64_t>& out, hipStream_t s) {
  uint64_t total = 0;
  for (auto& pc : pieces) total += pc.second;
  const uint64_t* p = gather_pinned(ctx, pieces, s);
  out.assign(p, p + total);
}

void tree_rows(const Tree& t, uint64_t index, std::vector<uint64_t>& out) {
  if (index >= t.height) throw ApiError(E_RANGE, "leaf index out of range");
  std::vector<std::pair<const uint64_t*, uint64_t>> pieces;
  for (auto& m : t.mats) {
    uint64_t row = index >> log2_exact(t.height / m.height);
    pieces.emplace_back(m.ptr + row * m.width, m.width);
  }
  std::vector<uint64_t> raw;
  gather_to_host(*t.ctx, pieces, raw, t.ctx->stream);
  out.clear();
  size_t off = 0;
  for (auto& m : t.mats) {
    out.insert(out.end(), raw.begin() + off, raw.begin() + off + m.width);
    out.resize(out.size() + (aligned_len(m.width, t.alignment) - m.width), 0);
    off += m.width;
  }
}


// prove_batch in two halves so that many trees can share one device gather:
// the pieces (rows, then missing siblings) in hint order, then the hints.
size_t prove_batch_pieces(const Tree& t, const std::vector<uint64_t>& indices,
                          std::vector<std::pair<const uint64_t*, uint64_t>>& pieces) {
  // unique indices, ascending (the BTreeSet of the reference's caller)
  std::vector<uint64_t> known(indices);
  std::sort(known.begin(), known.end());
  known.erase(std::unique(known.begin(), known.end()), known.end());
  if (!known.empty() && known.back() >= t.height) throw ApiError(E_RANGE, "query index out of range");
  // 1. rows of every matrix at every unique index, in leaf order
  const size_t first = pieces.size();
  pieces.reserve(first + known.size() * (t.mats.size() + 1 + t.depth));
  std::vector<unsigned> shift(t.mats.size());
  for (size_t m = 0; m < t.mats.size(); m++) shift[m] = log2_exact(t.height / t.mats[m].height);
  for (uint64_t idx : known) {
    for (size_t m = 0; m < t.mats.size(); m++) {
      const DevMat& mm = t.mats[m];
      pieces.emplace_back(mm.ptr + (idx >> shift[m]) * mm.width, mm.width);
    }
    if (t.salt.ptr) pieces.emplace_back(t.salt.ptr + idx * t.salt.width, t.salt.width);  // hinted after the rows
  }
  const size_t nrows = pieces.size() - first;
  // 2. missing siblings, level by level, left to right: in the sorted list a
  // node's sibling is known exactly when it is the neighbouring entry
  std::vector<uint64_t> parents;
  parents.reserve(known.size());
  for (unsigned layer = 0; layer < t.depth; layer++) {
    parents.clear();
    for (size_t q = 0; q < known.size(); q++) {
      const uint64_t node = known[q], parent = node / 2;
      if (q + 1 < known.size() && known[q + 1] == (node ^ 1)) {
        parents.push_back(parent);  // both children known
        q++;
        continue;
      }
      pieces.emplace_back(t.digests + (t.layer_off[layer] + (node ^ 1)) * 4, 4);
      parents.push_back(parent);
    }
    known.swap(parents);
  }
  return nrows;
}

// Emits the hints of pieces [first, last) (the first nrows are matrix rows)
// from their gathered words starting at raw; returns the words consumed.
size_t prove_batch_emit(const Tree& t, const std::vector<std::pair<const uint64_t*, uint64_t>>& pieces, size_t first,
                        size_t last, size_t nrows, const uint64_t* raw, Channel& tx) {
  size_t off = 0;
  std::vector<uint64_t> padded;
  // per opened leaf: one row per matrix (padded to the alignment), then the salt as it is
  const size_t per = t.mats.size() + (t.salt.ptr ? 1 : 0);
  for (size_t p = first; p < first + nrows; p++) {
    uint64_t w = pieces[p].second;
    const bool is_salt = t.salt.ptr && (p - first) % per == per - 1;
    uint64_t aw = is_salt ? w : aligned_len(w, t.alignment);
    if (aw == w) {
      tx.hint_fields(raw + off, w);  // no padding: straight from the gathered words
    } else {
      padded.assign(raw + off, raw + off + w);
      padded.resize(aw, 0);
      tx.hint_fields(padded.data(), aw);
    }
    off += w;
  }
  for (size_t p = first + nrows; p < last; p++) {
    tx.hint_commitment(raw + off);
    off += 4;
  }
  return off;
}

void tree_prove_batch(const Tree& t, const std::vector<uint64_t>& indices, Channel& tx, hipStream_t s) {
  std::vector<std::pair<const uint64_t*, uint64_t>> pieces;
  const size_t nrows = prove_batch_pieces(t, indices, pieces);
  const uint64_t* raw = gather_pinned(*t.ctx, pieces, s);
  prove_batch_emit(t, pieces, 0, pieces.size(), nrows, raw, tx);
}

void host_sponge_absorb(uint32_t hash, uint64_t* state, const uint64_t* in, size_t len) {
  const size_t rate = hash_rate(hash);
  size_t j = 0;
  while (j < len) {
    size_t rem = len - j;
    for (size_t k = 0; k < rate; k++) state[k] = k < rem ? in[j + k] : 0;
    if (hash == HASH_P2) p2_permute_fast(state, &host_p2_consts());
    else keccak_f(state);
    j += rate < rem ? rate : rem;
  }
}

void host_compress(uint32_t hash, const uint64_t* l, const uint64_t* r, uint64_t* out) {
  uint64_t s[25] = {};
  memcpy(s, l, 32);
  memcpy(s + 4, r, 32);
  if (hash == HASH_P2) p2_permute_fast(s, &host_p2_consts());
  else keccak_f(s);
  memcpy(out, s, 32);
}

Tree* commit_traces(Ctx& ctx, uint32_t hash, const std::vector<DevMat>& traces, unsigned log_blowup, hipStream_t s) {
  if (traces.empty()) throw ApiError(E_EMPTY, "at least one trace required");
  for (size_t i = 1; i < traces.size(); i++)
    if (traces[i].height < traces[i - 1].height) throw ApiError(E_UNSORTED, "traces must be sorted by height");
  for (auto& tr : traces)
    if (tr.height == 0 || (tr.height & (tr.height - 1))) throw ApiError(E_NOT_POW2, "trace height must be a power of two");
  const unsigned lmax = log2_exact(traces.back().height);
  std::vector<DevMat> ldes;
  try {
    for (auto& tr : traces) {
      unsigned lt = log2_exact(tr.height);
      uint64_t shift = GENERATOR;
      for (unsigned k = lt; k < lmax; k++) shift = gl_mul(shift, shift);
      DevMat m;
      m.height = tr.height << log_blowup;
      m.width = tr.width;
      m.owned = true;
      m.ptr = (uint64_t*)ctx.alloc(m.height * m.width * sizeof(uint64_t));
      ldes.push_back(m);
      lde_coset_batch(ctx, tr.ptr, tr.height, tr.width, log_blowup, shift, m.ptr, s);
    }
    ctx.check_cancel();
    return lmcs_build(ctx, hash, ldes, hash_alignment(hash), s);
  } catch (...) {
    for (auto& m : ldes) ctx.release(m.ptr);
    throw;
  }
}

// commit_quotient (p3-miden-lifted-stark/src/prover/quotient.rs:133-238).  The
// reference runs a plain iDFT over H on each chunk column t, scales coefficient
// k by (omega_J^-k)^t, zero-pads and runs a plain DFT over K.  Column t of the
// input holds q_t(g omega_J^t X) on H, so the result, q_t(g omega_K^i) for row
// bitrev(i), is the coset LDE of that column on the shift omega_J^-t: one
// inverse transform, a per-column twist of the coefficients (both base
// coordinates of a chunk share t) and a shift-1 forward transform, in the
// same passes as the trace LDE.  The single matrix is committed aligned.
Tree* commit_quotient(Ctx& ctx, uint32_t hash, const uint64_t* d_q, uint64_t n, uint64_t d, unsigned log_blowup,
                      hipStream_t s) {
  if (n == 0 || (n & (n - 1))) throw ApiError(E_NOT_POW2, "quotient trace height must be a power of two");
  if (d == 0 || (d & (d - 1))) throw ApiError(E_NOT_POW2, "constraint degree must be a power of two");
  if (d > (1ull << log_blowup)) throw ApiError(E_INVAL, "blowup must be at least the constraint degree");
  const unsigned logn = log2_exact(n), logd = log2_exact(d);
  if (logn + log_blowup > 32) throw ApiError(E_RANGE, "LDE height exceeds the field two-adicity");
  DevMat m;
  m.height = n << log_blowup;
  m.width = 2 * d;
  m.owned = true;
  m.ptr = (uint64_t*)ctx.alloc(m.height * m.width * sizeof(uint64_t));
  try {
    const ColTwist tw{gl_inv(gl_two_adic_generator(logn + logd)), logn + logd, 2};
    lde_coset_batch(ctx, d_q, n, 2 * d, log_blowup, 1, m.ptr, s, &tw);
    return lmcs_build(ctx, hash, {m}, hash_alignment(hash), s);
  } catch (...) {
    ctx.release(m.ptr);
    throw;
  }
}

}  // namespace p3g
