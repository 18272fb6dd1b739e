// MFMA "NT" GEMM for gfx950:  C[m, n] = sum_k A[m, k] * B[n, k]
//
// Both operands hold K contiguously in 64-byte stages, so every lane feeds
// the matrix core from one 16-byte LDS read per operand and K-chunk:
//   * bf16: v_mfma_f32_32x32x16_bf16, the 16 bytes are 8 consecutive k;
//   * fp32: v_mfma_f32_32x32x2_f32 issued 4 times, element s of the 16 bytes
//     feeds issue s.  Lanes 0-31 and 32-63 read the two 16-byte halves of a
//     32-byte K-chunk, so A and B see the same k <-> (lane, element)
//     assignment and the contraction covers each k exactly once.
//
// Operand addressing: byte address of (row r, stage s, chunk c) is
//   base + r * ld + s * stage + 16 c.
// Row-major operands with a short K (positions, X rows) use ld = K bytes,
// stage = 64.  Operands with a long K (the residuals R^T and X^T, K = data
// rows) are stored K-blocked: ld = 64, stage = rows * 64, so the tile a stage
// reads is one contiguous block (a row-major [1e3, 1e6] layout would put
// every row of a tile on its own 2 MB page and thrash the TLB).
//
// Work decomposition: a workgroup of WAVES_M x WAVES_N wave64s owns a
// BM x BN output tile (each wave WM x WN blocks of 32 x 32) and a range of
// K stages (split-K).  Stages are double-buffered in LDS (row stride 80
// bytes: the lanes served per LDS cycle by a ds_read_b128 hit disjoint
// banks), and the global loads run two stages ahead in registers.
//
// Tile order is XCD-aware: consecutive workgroup ids go round-robin to the
// 8 XCDs, so ids are mapped such that all workgroups that read the same B
// tile (mode 1) or the same K split (mode 2) land on one XCD and share its L2.
#pragma once

#include <type_traits>
#include <utility>

#include "common.h"

namespace edhmc {

constexpr int kStageBytes = 64;
#ifndef EDHMC_LDS_ROW_STRIDE
#define EDHMC_LDS_ROW_STRIDE 80
#endif
constexpr int kRowStride = EDHMC_LDS_ROW_STRIDE;   // LDS bytes per staged row (64 + pad)
static_assert(kRowStride >= kStageBytes && kRowStride % 16 == 0, "LDS row stride");

struct GemmShape {
  const char* A;      // [planes][M][K]
  int64_t lda_b;      // bytes between rows of A
  int64_t a_stage_b;  // bytes between K stages of A
  int64_t a_plane_b;  // bytes between planes of A
  const char* B;      // [N][K]
  int64_t ldb_b;
  int64_t b_stage_b;
  int m_tiles, n_tiles;
  int k_stages;          // K bytes / 64
  int ksplit;            // number of K ranges
  int stages_per_split;
  int swz;               // 0: plain, 1: group by n tile, 2: group by K split,
                         // 3: XCD owns a slice of M tiles (m_tiles % 8 == 0)
};

__device__ inline bool map_tile(const GemmShape& g, int id, int& tm, int& tn, int& ks) {
  if (g.swz == 1) {
    int xcd = id & 7, local = id >> 3;
    int grp = local / g.m_tiles;
    tm = local - grp * g.m_tiles;
    tn = xcd + 8 * grp;
    ks = 0;
    return tn < g.n_tiles;
  }
  if (g.swz == 3) {
    // each XCD owns m_tiles / 8 consecutive M tiles (its slice of A stays in
    // its L2) and walks the N tiles; m_tiles % 8 == 0
    const int xcd = id & 7, local = id >> 3, mper = g.m_tiles >> 3;
    const int grp = local / mper;
    tm = xcd * mper + (local - grp * mper);
    tn = grp;
    ks = 0;
    return tn < g.n_tiles;
  }
  int tiles = g.m_tiles * g.n_tiles;
  int t;
  if (g.swz == 2) {
    int xcd = id & 7, local = id >> 3;
    int grp = local / tiles;
    t = local - grp * tiles;
    ks = xcd + 8 * grp;
    if (ks >= g.ksplit) return false;
  } else {
    ks = id / tiles;
    t = id - ks * tiles;
  }
  tm = t % g.m_tiles;
  tn = t / g.m_tiles;
  return true;
}

inline int grid_size(const GemmShape& g) {
  if (g.swz == 1) return g.m_tiles * ((g.n_tiles + 7) / 8) * 8;
  if (g.swz == 2) return g.m_tiles * g.n_tiles * ((g.ksplit + 7) / 8) * 8;
  return g.m_tiles * g.n_tiles * g.ksplit;
}

// The K loop addresses a chunk as stage base + 32-bit lane offset: the
// largest offset inside one stage tile must fit (A planes, rows of A and B).
template <typename Cfg, int A_PLANES>
inline bool gemm_offsets_fit(const GemmShape& g) {
  const int64_t a = (int64_t)(A_PLANES - 1) * g.a_plane_b + (int64_t)Cfg::BM * g.lda_b;
  const int64_t b = (int64_t)Cfg::BN * g.ldb_b;
  return a < (int64_t)UINT32_MAX && b < (int64_t)UINT32_MAX;
}

#ifndef EDHMC_PREFETCH_DEPTH
#define EDHMC_PREFETCH_DEPTH 3
#endif
constexpr int kPrefetchDepth = EDHMC_PREFETCH_DEPTH;   // register stages in flight (2 or 3)
#ifndef EDHMC_MFMA_PRIO
#define EDHMC_MFMA_PRIO 0
#endif
constexpr bool kMfmaPrio = EDHMC_MFMA_PRIO != 0;
static_assert(kPrefetchDepth == 2 || kPrefetchDepth == 3, "prefetch depth");

// Word 3 of a raw (unswizzled, stride 0) buffer resource on gfx9-family
// CDNA parts: 32-bit data format, no index, offsets checked against the
// byte range in word 2.
constexpr int kBufferRsrcWord3 = 0x00020000;

template <typename T> struct Mfma;

template <> struct Mfma<bf16_t> {
  __device__ static inline void chunk(f32x16_t& acc, const uint4& a, const uint4& b) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                  __builtin_bit_cast(bf16x8_t, b),
                                                  acc, 0, 0, 0);
  }
};

template <> struct Mfma<float> {
  __device__ static inline void chunk(f32x16_t& acc, const uint4& a, const uint4& b) {
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__builtin_bit_cast(float, a.x),
                                               __builtin_bit_cast(float, b.x), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__builtin_bit_cast(float, a.y),
                                               __builtin_bit_cast(float, b.y), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__builtin_bit_cast(float, a.z),
                                               __builtin_bit_cast(float, b.z), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__builtin_bit_cast(float, a.w),
                                               __builtin_bit_cast(float, b.w), acc, 0, 0, 0);
  }
};

// Row of 32x32 accumulator element r held by `lane` (column is lane & 31).
__device__ inline int acc_row(int r, int lane) {
  return ((r >> 2) << 3) + ((lane >> 5) << 2) + (r & 3);
}

// WM x WN blocks of 32 x 32 per wave, WAVES_M x WAVES_N waves, KSB bytes of
// K per pipeline stage (64 or 128; the operand rows are staged in LDS at
// KSB + 16 bytes per row).
template <int WM, int WN, int WAVES_M, int WAVES_N, int KSB = kStageBytes>
struct TileCfg {
  static constexpr int kWM = WM, kWN = WN, kWavesM = WAVES_M, kWavesN = WAVES_N;
  static constexpr int kThreads = 64 * WAVES_M * WAVES_N;
  static constexpr int BM = 32 * WM * WAVES_M;
  static constexpr int BN = 32 * WN * WAVES_N;
  static constexpr int KS = KSB;
  static constexpr int RS = KSB + (kRowStride - kStageBytes);
  static_assert(KSB % 64 == 0, "stage bytes");
};

// Epilogue context handed to the epilogue functor.
struct EpiPos {
  int m0, n0;     // global row / column of this wave's first accumulator block
  int lane, wm, wn, tm, tn, ks;
  char* smem;     // free LDS (>= the main loop's footprint) after a barrier
  int smem_bytes; // its size
};

#ifndef EDHMC_GEMM_WPE
#define EDHMC_GEMM_WPE 2
#endif

// LDS bytes an epilogue asks for (its static lds_bytes<Cfg>(), 0 without
// one); the kernel reserves the larger of that and its main loop's buffers.
template <typename Epi, typename Cfg, typename = void>
struct EpiLds {
  static constexpr int value = 0;
};
template <typename Epi, typename Cfg>
struct EpiLds<Epi, Cfg, std::void_t<decltype(Epi::template lds_bytes<Cfg>())>> {
  static constexpr int value = Epi::template lds_bytes<Cfg>();
};
// Epilogue operands loaded before the K loop: an epilogue with a
// prefetch<Cfg>(pos) member gets its result as apply's third argument, so
// the loads' latency hides behind the main loop (an epilogue that reads
// per-element state, e.g. the leapfrog's positions and momenta).
struct NoPre {};
template <typename Epi, typename Cfg, typename = void>
struct EpiPre {
  __device__ static NoPre load(const Epi&, const EpiPos&) { return NoPre{}; }
  template <typename Acc>
  __device__ static void apply(const Epi& e, Acc& acc, const EpiPos& p, const NoPre&) {
    e.template apply<Cfg>(acc, p);
  }
};
template <typename Epi, typename Cfg>
struct EpiPre<Epi, Cfg, std::void_t<decltype(std::declval<const Epi&>().template prefetch<Cfg>(
                            std::declval<const EpiPos&>()))>> {
  __device__ static auto load(const Epi& e, const EpiPos& p) { return e.template prefetch<Cfg>(p); }
  template <typename Acc, typename Pre>
  __device__ static void apply(const Epi& e, Acc& acc, const EpiPos& p, const Pre& q) {
    e.template apply<Cfg>(acc, p, q);
  }
};

template <typename T, typename Cfg, int A_PLANES, typename Epi>
__global__ void __launch_bounds__(Cfg::kThreads) __attribute__((amdgpu_waves_per_eu(1, EDHMC_GEMM_WPE)))
gemm_nt_kernel(GemmShape g, Epi epi) {
  constexpr int NT = Cfg::kThreads;
  constexpr int WM = Cfg::kWM, WN = Cfg::kWN, WAVES_M = Cfg::kWavesM;
  constexpr int BM = Cfg::BM, BN = Cfg::BN;
  constexpr int AR = BM * A_PLANES;
  constexpr int ROWS = AR + BN;
  constexpr int CPR = Cfg::KS / 16;              // 16-byte chunks per row
  constexpr int CHUNKS = ROWS * CPR;
  constexpr int PER_T = (CHUNKS + NT - 1) / NT;
  constexpr int BUF = ROWS * Cfg::RS;
  constexpr int SMEM = 2 * BUF > EpiLds<Epi, Cfg>::value ? 2 * BUF : EpiLds<Epi, Cfg>::value;
  __shared__ __align__(16) char smem[SMEM];

  int tm, tn, ks;
  if (!map_tile(g, blockIdx.x, tm, tn, ks)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;
  const int s_begin = ks * g.stages_per_split;
  const int s_end = min(g.k_stages, s_begin + g.stages_per_split);

  // Chunk q of a stage (16 bytes) is row (tid + q NT) / CPR, column
  // (tid + q NT) % CPR of the combined [A planes | B] tile.  Whether a chunk
  // is in A or B depends only on q (AR is a multiple of NT / CPR), so the
  // address selection below resolves at compile time after unrolling.
  static_assert((AR * CPR) % NT == 0 || CHUNKS <= NT, "A/B split must fall on a chunk round");
  // When every round of NT chunks lies wholly in A or in B, a chunk's
  // address is a wave-uniform stage base (scalar registers) plus a per-lane
  // 32-bit offset fixed for the whole K loop: no per-stage vector address
  // arithmetic and no 64-bit address registers (which the allocator would
  // otherwise share with the prefetch registers, serialising loads behind
  // the LDS stores).  gemm_offsets_fit() checks the offsets fit in 32 bits.
  constexpr bool kRoundSplit = (AR * CPR) % NT == 0;
  const char* a_base = g.A + (int64_t)(tm * BM) * g.lda_b;
  const char* b_base = g.B + (int64_t)(tn * BN) * g.ldb_b;
  uint32_t loff[PER_T];
#pragma unroll
  for (int q = 0; q < PER_T; ++q) {
    int c = tid + q * NT;
    if (c >= CHUNKS) c = CHUNKS - 1;  // duplicate the last chunk (same data)
    const int row = c / CPR, col = c - row * CPR;
    if (row < AR) {
      const int p = row / BM, r = row - p * BM;
      loff[q] = (uint32_t)(p * g.a_plane_b + (int64_t)r * g.lda_b + col * 16);
    } else {
      loff[q] = (uint32_t)((int64_t)(row - AR) * g.ldb_b + col * 16);
    }
  }
  auto chunk_src = [&](int q, int s) -> const char* {
    if (kRoundSplit) {
      const bool in_a = q * NT < AR * CPR;   // compile-time after unrolling
      const char* base = in_a ? a_base + (int64_t)s * g.a_stage_b : b_base + (int64_t)s * g.b_stage_b;
      return base + loff[q];
    }
    int c = tid + q * NT;
    if (c >= CHUNKS) c = CHUNKS - 1;
    const int row = c / CPR;
    return (row < AR ? a_base + (int64_t)s * g.a_stage_b : b_base + (int64_t)s * g.b_stage_b) +
           loff[q];
  };
  // Per-stage buffer resources (scalar registers): base = the stage's tile
  // origin in A and in B, range = the tile's extent, so a chunk load is one
  // buffer_load with a lane offset fixed for the K loop and no vector
  // address arithmetic.
  typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
  struct stage_rsrc { __amdgpu_buffer_rsrc_t a, b; };
  const int a_extent = (int)((A_PLANES - 1) * g.a_plane_b + (int64_t)BM * g.lda_b);
  const int b_extent = (int)((int64_t)BN * g.ldb_b);
  auto make_stage_rsrc = [&](int s) -> stage_rsrc {
    stage_rsrc r;
    r.a = __builtin_amdgcn_make_buffer_rsrc((void*)(a_base + (int64_t)s * g.a_stage_b), (short)0,
                                            a_extent, kBufferRsrcWord3);
    r.b = __builtin_amdgcn_make_buffer_rsrc((void*)(b_base + (int64_t)s * g.b_stage_b), (short)0,
                                            b_extent, kBufferRsrcWord3);
    return r;
  };
  auto load_chunk = [&](const stage_rsrc& r, int q, int s) -> u32x4_t {
    if (kRoundSplit) {
      const bool in_a = q * NT < AR * CPR;   // compile-time after unrolling
      return __builtin_amdgcn_raw_buffer_load_b128(in_a ? r.a : r.b, (int)loff[q], 0, 0);
    }
    const uint4 x = *(const uint4*)chunk_src(q, s);
    u32x4_t v;
    v.x = x.x; v.y = x.y; v.z = x.z; v.w = x.w;
    return v;
  };
  auto chunk_dst = [&](int q) -> int {
    int c = tid + q * NT;
    if (c >= CHUNKS) c = CHUNKS - 1;
    const int row = c / CPR, col = c - row * CPR;
    return row * Cfg::RS + col * 16;
  };

  EpiPos pos;
  pos.m0 = tm * BM + wm * 32 * WM;
  pos.n0 = tn * BN + wn * 32 * WN;
  pos.lane = lane; pos.wm = wm; pos.wn = wn; pos.tm = tm; pos.tn = tn; pos.ks = ks;
  pos.smem = smem;
  pos.smem_bytes = SMEM;
  const auto pre = EpiPre<Epi, Cfg>::load(epi, pos);

  f32x16_t acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  const int lrow = lane & 31, lhalf = (lane >> 5) * 16;
  auto compute = [&](int buf) {
    const char* base = smem + buf * BUF;
    // a wave in its MFMA phase issues ahead of waves loading or storing
    if (kMfmaPrio) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int kc = 0; kc < Cfg::KS / 32; ++kc) {
      uint4 a[A_PLANES][WM], b[WN];
#pragma unroll
      for (int p = 0; p < A_PLANES; ++p)
#pragma unroll
        for (int i = 0; i < WM; ++i)
          a[p][i] = *(const uint4*)(base + (p * BM + wm * 32 * WM + 32 * i + lrow) * Cfg::RS +
                                    kc * 32 + lhalf);
#pragma unroll
      for (int j = 0; j < WN; ++j)
        b[j] = *(const uint4*)(base + (AR + wn * 32 * WN + 32 * j + lrow) * Cfg::RS +
                               kc * 32 + lhalf);
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
          for (int p = 0; p < A_PLANES; ++p) Mfma<T>::chunk(acc[i][j], a[p][i], b[j]);
    }
    if (kMfmaPrio) __builtin_amdgcn_s_setprio(0);
  };

  // Register prefetch sets: each is one vector-register tuple (ext_vector of
  // 4 PER_T dwords), never an array, so it stays in VGPRs.  With depth D the
  // loads of stage s + D are issued while stage s is computed and stage s + 1
  // is stored to LDS, so D - 1 stages of compute hide each load's latency.
  typedef uint32_t PreT __attribute__((ext_vector_type(4 * PER_T)));
  PreT r0, r1, r2;
#define EDHMC_GLOAD(PRE, S)                                              \
  {                                                                      \
    const stage_rsrc rs_ = make_stage_rsrc(S);                           \
    _Pragma("unroll") for (int q = 0; q < PER_T; ++q) {                  \
      const u32x4_t x_ = load_chunk(rs_, q, (S));                        \
      PRE[4 * q] = x_.x; PRE[4 * q + 1] = x_.y;                          \
      PRE[4 * q + 2] = x_.z; PRE[4 * q + 3] = x_.w;                      \
    }                                                                    \
    __builtin_amdgcn_sched_barrier(0);                                   \
  }
#define EDHMC_SSTORE(PRE, BUFI)                                          \
  _Pragma("unroll") for (int q = 0; q < PER_T; ++q) {                    \
    uint4 x_;                                                            \
    x_.x = PRE[4 * q]; x_.y = PRE[4 * q + 1];                            \
    x_.z = PRE[4 * q + 2]; x_.w = PRE[4 * q + 3];                        \
    *(uint4*)(smem + (BUFI) * BUF + chunk_dst(q)) = x_;                  \
  }
// One K stage: optionally load stage S + DEPTH into NXT, compute from LDS
// buffer (S - s_begin) & 1, optionally store the stage held in CUR (S + 1)
// into the other buffer, barrier.  In the steady-state loops the load is
// never under a runtime branch: a load under a branch makes the compiler's
// wait-count merge at the join assume the newest loads may be the ones the
// store needs, i.e. wait for every outstanding load, which would turn the
// prefetch into a stall per stage.  Only the peeled last stages branch.
#define EDHMC_STEP(S, CUR, NXT, DEPTH, LOAD, STORE)                      \
  {                                                                      \
    const int buf_ = ((S) - s_begin) & 1;                                \
    if (LOAD) { EDHMC_GLOAD(NXT, (S) + (DEPTH)) }                        \
    compute(buf_);                                                       \
    if (STORE) { EDHMC_SSTORE(CUR, buf_ ^ 1) }                           \
    __syncthreads();                                                     \
  }
  if constexpr (kPrefetchDepth == 3) {
    // invariant at the top of step s: stage s in LDS, stages s+1, s+2 in
    // registers; the loop body rotates r1 (s+1), r2 (s+2), r0 (free)
    if (s_begin < s_end) {
      EDHMC_GLOAD(r0, s_begin)
      EDHMC_SSTORE(r0, 0)
      if (s_begin + 1 < s_end) { EDHMC_GLOAD(r1, s_begin + 1) }
      if (s_begin + 2 < s_end) { EDHMC_GLOAD(r2, s_begin + 2) }
    }
    __syncthreads();
    int s = s_begin;
    for (; s + 5 < s_end; s += 3) {
      EDHMC_STEP(s, r1, r0, 3, true, true)
      EDHMC_STEP(s + 1, r2, r1, 3, true, true)
      EDHMC_STEP(s + 2, r0, r2, 3, true, true)
    }
    // at most five stages left; the rotation is the loop's (s = s_begin mod 3)
    if (s < s_end) EDHMC_STEP(s, r1, r0, 3, s + 3 < s_end, s + 1 < s_end)
    if (s + 1 < s_end) EDHMC_STEP(s + 1, r2, r1, 3, s + 4 < s_end, s + 2 < s_end)
    if (s + 2 < s_end) EDHMC_STEP(s + 2, r0, r2, 3, false, s + 3 < s_end)
    if (s + 3 < s_end) EDHMC_STEP(s + 3, r1, r0, 3, false, s + 4 < s_end)
    if (s + 4 < s_end) EDHMC_STEP(s + 4, r2, r1, 3, false, false)
  } else {
    // depth 2: stage s in LDS, stage s + 1 in r1 at the top of step s
    if (s_begin < s_end) {
      EDHMC_GLOAD(r0, s_begin)
      EDHMC_SSTORE(r0, 0)
      if (s_begin + 1 < s_end) { EDHMC_GLOAD(r1, s_begin + 1) }
    }
    __syncthreads();
    int s = s_begin;
    for (; s + 3 < s_end; s += 2) {
      EDHMC_STEP(s, r1, r0, 2, true, true)
      EDHMC_STEP(s + 1, r0, r1, 2, true, true)
    }
    if (s < s_end) EDHMC_STEP(s, r1, r0, 2, s + 2 < s_end, s + 1 < s_end)
    if (s + 1 < s_end) EDHMC_STEP(s + 1, r0, r1, 2, false, s + 2 < s_end)
    if (s + 2 < s_end) EDHMC_STEP(s + 2, r1, r0, 2, false, false)
  }
#undef EDHMC_STEP
#undef EDHMC_SSTORE
#undef EDHMC_GLOAD

  EpiPre<Epi, Cfg>::apply(epi, acc, pos, pre);
}

// ---------------------------------------------------------------------------
// Epilogue: store the fp32 tile (split-K partials: out[ks][m][n]).
// ---------------------------------------------------------------------------
struct EpiStoreF32 {
  float* out;
  int64_t ldc;        // elements between rows
  int64_t split_b;    // elements between split planes
  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p) const {
    float* o = out + (int64_t)p.ks * split_b;
#pragma unroll
    for (int i = 0; i < Cfg::kWM; ++i)
#pragma unroll
      for (int j = 0; j < Cfg::kWN; ++j) {
        const int n = p.n0 + 32 * j + (p.lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int m = p.m0 + 32 * i + acc_row(r, p.lane);
          o[(int64_t)m * ldc + n] = acc[i][j][r];
        }
      }
  }
};

}  // namespace edhmc
