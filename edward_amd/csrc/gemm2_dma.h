// GEMM2 of the bf16 gradient pass, G[c, d] = sum_n R[c, n] X[n, d], on
// K-blocked operands (R^T and X^T stored as 64-byte stages of 32 data rows:
// byte (row, stage s, chunk q) at base + s * rows * 64 + row * 64 + 16 q).
//
// One workgroup of four waves (one per SIMD) owns a 256 x 256 output tile
// and a range of K stages (split-K); each wave holds 128 x 128 of it in
// sixteen 32 x 32 accumulator blocks, so a 16-byte operand fragment read
// from LDS feeds four MFMAs.  Stages reach LDS by buffer loads that write
// LDS directly (no register staging, no LDS stores): each load instruction
// fills 1 KB of LDS, 16 bytes per lane at lane * 16, and the lanes fetch
// the chunk that belongs there.  Chunk q of operand row r sits in LDS slot
// q ^ ((r >> 2) & 3) of the row's 64 bytes, which makes the fragment reads
// (lanes 0-31: rows r0..r0+31 at one chunk) hit disjoint banks in every
// group of 16 lanes.  Five stages are in flight in a ring of five 32 KB LDS
// slots (160 KB): distinct arrays addressed with compile-time slot names,
// so the compiler can tell which slot a fragment read and a direct load
// touch and does not make the reads wait for every load in flight.  One
// barrier per stage, placed between the two 16-k halves of the stage, so
// the next stage's first fragments are read while the second half's MFMAs
// run.
//
// Per output element the accumulation is the gemm_nt kernel's: the same
// 32x32x16 MFMA, the same k <-> (lane, element) assignment, stages in
// increasing order within the split, the same split boundaries; the
// results are the same bits.
#pragma once

#include "common.h"
#include "gemm_nt.h"

namespace edhmc {

constexpr int kG2Tile = 256;                    // output tile (chains x columns)
constexpr int kG2Ring = 5;                      // stages in flight (LDS slots)
constexpr int kG2StageLds = 2 * kG2Tile * 64;   // LDS bytes of one stage (A | B)
constexpr int kG2Loads = kG2StageLds / 1024 / 4;   // 1 KB direct loads per wave and stage
static_assert(kG2Ring * kG2StageLds <= 160 * 1024, "LDS ring");

// s_waitcnt immediates (gfx9 encoding: vmcnt[3:0] bits 3:0, expcnt 6:4,
// lgkmcnt 11:8, vmcnt[5:4] bits 15:14): wait until at most n vector-memory
// operations of this wave are outstanding and every LDS operation is done.
constexpr int vmcnt_lgkm0(int n) { return 0x0070 | (n & 15) | ((n >> 4) << 14); }

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
gemm2_dma_kernel(GemmShape g, float* __restrict__ out, int64_t ldc, int64_t split_b) {
  constexpr int NS = kG2Ring, T = kG2Tile;
  __shared__ __align__(16) char slot0[kG2StageLds], slot1[kG2StageLds], slot2[kG2StageLds],
      slot3[kG2StageLds], slot4[kG2StageLds];
  typedef __attribute__((address_space(3))) void lds_void;

  int tm, tn, ks;
  if (!map_tile(g, blockIdx.x, tm, tn, ks)) return;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = w & 1, wn = w >> 1;
  const int s_begin = ks * g.stages_per_split;
  // stages of this split (the host launches the kernel only when every
  // split has at least NS)
  const int S = min(g.k_stages, s_begin + g.stages_per_split) - s_begin;

  // Loads: wave w fetches A rows 64 w + 16 i + (l >> 2) and B rows
  // 64 w + 16 i + (l >> 2), i < 4; lane l fills LDS slot l & 3 of its row,
  // i.e. it fetches chunk (l & 3) ^ ((l >> 4) & 3).
  const char* const a_base = g.A + (int64_t)tm * T * g.lda_b;
  const char* const b_base = g.B + (int64_t)tn * T * g.ldb_b;
  const int ld_off = w * (T / 4) * 64 + (lane >> 2) * 64 + (((lane & 3) ^ ((lane >> 4) & 3)) << 4);
#define EDHMC_G2_LOAD_STAGE(U, SLOT)                                               \
  {                                                                                \
    const __amdgpu_buffer_rsrc_t ra_ = __builtin_amdgcn_make_buffer_rsrc(          \
        (void*)(a_base + (int64_t)(s_begin + (U)) * g.a_stage_b), (short)0,        \
        T * 64, kBufferRsrcWord3);                                                 \
    const __amdgpu_buffer_rsrc_t rb_ = __builtin_amdgcn_make_buffer_rsrc(          \
        (void*)(b_base + (int64_t)(s_begin + (U)) * g.b_stage_b), (short)0,        \
        T * 64, kBufferRsrcWord3);                                                 \
    _Pragma("unroll") for (int i = 0; i < kG2Loads / 2; ++i)                       \
      __builtin_amdgcn_raw_ptr_buffer_load_lds(                                    \
          ra_, (lds_void*)((SLOT) + w * (T / 4) * 64 + i * 1024), 16,              \
          ld_off + i * 1024, 0, 0, 0);                                             \
    _Pragma("unroll") for (int i = 0; i < kG2Loads / 2; ++i)                       \
      __builtin_amdgcn_raw_ptr_buffer_load_lds(                                    \
          rb_, (lds_void*)((SLOT) + T * 64 + w * (T / 4) * 64 + i * 1024), 16,     \
          ld_off + i * 1024, 0, 0, 0);                                             \
  }

  // Fragment reads: block i of the wave's A rows (row wm * 128 + 32 i +
  // (l & 31)), 16-k half kc: chunk 2 kc + (l >> 5), at its swizzled slot;
  // (M) zeroes the fragments (stages past the split's end).  Macros, not
  // lambdas over array references: those put the fragment arrays in scratch.
  const int rsw = ((lane & 31) >> 2) & 3;
  const int a_row = (wm * 128 + (lane & 31)) * 64;
  const int b_row = T * 64 + (wn * 128 + (lane & 31)) * 64;
  const int off_kc0 = ((0 + (lane >> 5)) ^ rsw) << 4;
  const int off_kc1 = ((2 + (lane >> 5)) ^ rsw) << 4;
  uint4 a0[4], b0[4], a1[4], b1[4];
#define EDHMC_G2_READ(SLOT, OFF, A, B, M)                                          \
  {                                                                                \
    const char* base_ = (SLOT);                                                    \
    _Pragma("unroll") for (int i = 0; i < 4; ++i)                                  \
      A[i] = *(const uint4*)(base_ + a_row + i * 2048 + (OFF));                    \
    _Pragma("unroll") for (int j = 0; j < 4; ++j)                                  \
      B[j] = *(const uint4*)(base_ + b_row + j * 2048 + (OFF));                    \
    if ((M) != ~0u) {                                                              \
      _Pragma("unroll") for (int i = 0; i < 4; ++i) {                              \
        A[i].x &= (M); A[i].y &= (M); A[i].z &= (M); A[i].w &= (M);                \
        B[i].x &= (M); B[i].y &= (M); B[i].z &= (M); B[i].w &= (M);                \
      }                                                                            \
    }                                                                              \
  }
#define EDHMC_G2_MFMAS(A, B)                                                       \
  _Pragma("unroll") for (int i = 0; i < 4; ++i)                                    \
    _Pragma("unroll") for (int j = 0; j < 4; ++j) Mfma<bf16_t>::chunk(acc[i][j], A[i], B[j]);

  f32x16_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  auto barrier = [] {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };

  // prologue: stages 0 .. NS - 1 in flight, stage 0 complete everywhere
  EDHMC_G2_LOAD_STAGE(0, slot0)
  EDHMC_G2_LOAD_STAGE(1, slot1)
  EDHMC_G2_LOAD_STAGE(2, slot2)
  EDHMC_G2_LOAD_STAGE(3, slot3)
  EDHMC_G2_LOAD_STAGE(4, slot4)
  __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0(kG2Loads * (NS - 1)));
  barrier();
  EDHMC_G2_READ(slot0, off_kc0, a0, b0, ~0u)

  // Stage t (in slot CUR): its second half's fragments, its first half's
  // MFMAs; wait until this wave's loads of stage t + 1 are done (stages
  // t + 2 .. t + NS - 1 may still be in flight) and its reads of stage t
  // are done; barrier (stage t + 1 complete in LDS, nobody reads CUR any
  // more); stage t + NS into CUR (past the end: the last stage again, never
  // read as data, so every step has the same loads in flight); stage t + 1's
  // first fragments; the second half's MFMAs.  The scheduling barriers keep
  // this order.
#define EDHMC_G2_STEP(CUR, NXT, MC, MN)                                            \
  {                                                                                \
    EDHMC_G2_READ(CUR, off_kc1, a1, b1, MC)                                        \
    __builtin_amdgcn_sched_barrier(0);                                             \
    EDHMC_G2_MFMAS(a0, b0)                                                         \
    __builtin_amdgcn_sched_barrier(0);                                             \
    __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0(kG2Loads * (NS - 2)));                  \
    barrier();                                                                     \
    __builtin_amdgcn_sched_barrier(0);                                             \
    EDHMC_G2_LOAD_STAGE(min(t + NS, S - 1), CUR)                                   \
    __builtin_amdgcn_sched_barrier(0);                                             \
    EDHMC_G2_READ(NXT, off_kc0, a0, b0, MN)                                        \
    __builtin_amdgcn_sched_barrier(0);                                             \
    EDHMC_G2_MFMAS(a1, b1)                                                         \
    __builtin_amdgcn_sched_barrier(0);                                             \
    ++t;                                                                           \
  }
  // Groups of NS steps that all have a next stage, then the last group:
  // stages NS full .. NS full + NS - 1, those at or past S (their slots hold
  // copies of real stages) read as zeros, which leave the sums as they are.
  int t = 0;
  const int full = (S - 1) / NS;
  for (int it = 0; it < full; ++it) {
    EDHMC_G2_STEP(slot0, slot1, ~0u, ~0u)
    EDHMC_G2_STEP(slot1, slot2, ~0u, ~0u)
    EDHMC_G2_STEP(slot2, slot3, ~0u, ~0u)
    EDHMC_G2_STEP(slot3, slot4, ~0u, ~0u)
    EDHMC_G2_STEP(slot4, slot0, ~0u, ~0u)
  }
  {
    const int t_last = NS * full;
    const uint32_t m1 = t_last + 1 < S ? ~0u : 0u;
    const uint32_t m2 = t_last + 2 < S ? ~0u : 0u;
    const uint32_t m3 = t_last + 3 < S ? ~0u : 0u;
    const uint32_t m4 = t_last + 4 < S ? ~0u : 0u;
    EDHMC_G2_STEP(slot0, slot1, ~0u, m1)
    EDHMC_G2_STEP(slot1, slot2, m1, m2)
    EDHMC_G2_STEP(slot2, slot3, m2, m3)
    EDHMC_G2_STEP(slot3, slot4, m3, m4)
    EDHMC_G2_READ(slot4, off_kc1, a1, b1, m4)
    EDHMC_G2_MFMAS(a0, b0)
    EDHMC_G2_MFMAS(a1, b1)
  }
  // no direct load may still write this workgroup's LDS when it ends
  __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0(0));
#undef EDHMC_G2_STEP
#undef EDHMC_G2_LOAD_STAGE
#undef EDHMC_G2_READ
#undef EDHMC_G2_MFMAS

  float* o = out + (int64_t)ks * split_b;
  const int m0 = tm * T + wm * 128, c0 = tn * T + wn * 128;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = c0 + 32 * j + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r)
        o[(int64_t)(m0 + 32 * i + acc_row(r, lane)) * ldc + n] = acc[i][j][r];
    }
}

}  // namespace edhmc
