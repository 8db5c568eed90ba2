// Reduced-precision grouped GEMMs for the effective-Hamiltonian matvec
// (real data; complex data in the split-BF16 modes).
//
// Same task / segment / tile lists as gemm_f64.hip.  A operands are FP32 in
// memory (the plan keeps FP32 copies of the environment stacks, and the
// intermediates are FP32); B operands are FP64 (the Krylov vector) or FP32;
// the output is FP32 (intermediate) or FP64 (the Krylov output, written
// through a row-address table).  Accumulation is FP32 in the MFMA units.
//
//  * gemm_f32_kernel: v_mfma_f32_16x16x4_f32 on FP32 operands.
//  * gemm_bf16_kernel: split-BF16.  Each FP32 operand value x is split into
//    bf16 pieces h = bf16(x), m = bf16(x - h), l = bf16(x - h - m) (each
//    subtraction exact in FP32) and the product is formed from the piece
//    products on v_mfma_f32_16x16x32_bf16:
//      NPC = 3:  h.h + h.l' + l'.h  with two pieces (l' = bf16(x - h)),
//                ~16 significant bits per operand;
//      NPC = 6:  h.h + h.m + m.h + h.l + l.h + m.m  with three pieces,
//                ~24 bits (FP32-like).
//    In the 6-term mode, stage 3's real operands are split once, when their K
//    slice is staged, and LDS holds one bf16 plane per piece; otherwise
//    operands are staged as FP32 and split when the MFMA fragments are read.
//
// MFMA layouts (measured on gfx950 with an in-bounds probe):
//   f32 16x16x4:   A lane l: A[l % 16][l / 16];  B lane l: B[l / 16][l % 16]
//   bf16 16x16x32: A lane l: A[l % 16][8 (l / 16) + j], j < 8;
//                  B lane l: B[8 (l / 16) + j][l % 16]
//   both:          D lane l, register r: D[4 (l / 16) + r][l % 16]
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "gemm.h"
#include "tile_io.h"

namespace qlg {

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {

__device__ inline float load_f32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}
template <typename T>
__device__ inline float load_as_f32(__amdgpu_buffer_rsrc_t r, uint32_t elem) {
  if constexpr (sizeof(T) == 8)
    return (float)load_f64(r, elem == kOob ? kOob : elem * 8u);
  else
    return load_f32(r, elem == kOob ? kOob : elem * 4u);
}

// output row addresses (elements from the task's C base): table or stride
template <int FM>
__device__ inline void out_rows(const GemmTask &tk, const GemmTile &tl, const Slots &slots, int wm0, int lane,
                                int64_t (&crow)[FM][4]) {
  if (tk.c_rowtab) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = tl.m0 + wm0 + i * 16 + 4 * (lane >> 4) + r;
        crow[i][r] = slots.crows[tk.c_row + min(row, tk.M - 1)] - tk.c_off;
      }
  } else {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) crow[i][r] = (int64_t)(tl.m0 + wm0 + i * 16 + 4 * (lane >> 4) + r) * tk.ldc;
  }
}

template <int FM, int FN, typename TC>
__device__ inline void store_tile(const GemmTask &tk, const GemmTile &tl, const Slots &slots, int wm0, int wn0,
                                  int lane, const flt4 (&acc)[FM][FN]) {
  int64_t crow[FM][4];
  out_rows<FM>(tk, tl, slots, wm0, lane, crow);
  TC *C = reinterpret_cast<TC *>(slots.p[tk.c_slot]) + tk.c_off;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = tl.n0 + wn0 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = tl.m0 + wm0 + i * 16 + 4 * (lane >> 4) + r;
        if (row < tk.M && col < tk.N) C[crow[i][r] + col] = (TC)acc[i][j][r];
      }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// FP32 MFMA.  LDS: As[k][m], Bs[k][n] (FP32), BK-deep slices, ping-pong.
template <int BM, int BN, int BK, typename TB, typename TC>
__global__ __launch_bounds__(256) void gemm_f32_kernel(const GemmTile *__restrict__ tiles,
                                                       const GemmTask *__restrict__ tasks,
                                                       const GemmSeg *__restrict__ segs, Slots slots, int ntiles) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int A_PER = BM * BK / 256, B_PER = BK * BN / 256;
  __shared__ float As[2][BK][BM + 1];
  __shared__ float Bs[2][BK][BN + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  if ((int)blockIdx.x >= ntiles) return;
  if (slots.skip && *slots.skip) return;   // the Krylov solve has stopped
  const GemmTile tl = tiles[blockIdx.x];
  if (tl.task < 0) return;
  const GemmTask tk = tasks[tl.task];
  const int M = tk.M, N = tk.N;

  flt4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
  float ra[A_PER], rb[B_PER];

  auto b_kn = [&](int i, bool tb, int &k, int &n) {
    const int e = tid + i * 256;
    n = tb ? e / BK : e % BN;
    k = tb ? e % BK : e / BN;
  };
  // Per-segment addressing (set up at each segment switch).  Since 256 is a
  // multiple of the slice depth and of BN, element i of this thread sits at
  // an affine position: offset off0 + i * soff (plus the slice's k0 term),
  // k = kt + i * sk and row (column) r0 + i * sr, with the steps uniform.
  __amdgpu_buffer_rsrc_t ar, br;
  uint32_t a_off0 = 0, b_off0 = 0, a_soff = 0, b_soff = 0, bstep = 1;
  int a_kt = 0, b_kt = 0, b_sk = 0, a_r0 = 0, b_r0 = 0, b_sr = 0, segK = 0;
  auto setup = [&](const GemmSeg &sg) {
    ar = make_rsrc(reinterpret_cast<const float *>(slots.p[sg.a_slot]) + sg.a_off);
    br = make_rsrc(reinterpret_cast<const TB *>(slots.p[sg.b_slot]) + sg.b_off);
    const uint32_t lda = (uint32_t)sg.lda, ldb = (uint32_t)sg.ldb;
    const bool tb = sg.flags & SEG_TRANS_B;
    segK = sg.K;
    // A element e = tid + 256 i: (m, k) = (e / BK, e % BK)
    a_r0 = tl.m0 + tid / BK;
    a_kt = tid % BK;
    a_off0 = (uint32_t)a_r0 * lda + (uint32_t)a_kt;
    a_soff = (uint32_t)(256 / BK) * lda;
    // B element e: (k, n) = (e / BN, e % BN), or (e % BK, e / BK) when transposed
    const int bn = tb ? tid / BK : tid % BN, bk = tb ? tid % BK : tid / BN;
    b_r0 = tl.n0 + bn;
    b_kt = bk;
    b_sr = tb ? 256 / BK : 0;
    b_sk = tb ? 0 : 256 / BN;
    b_off0 = tb ? (uint32_t)b_r0 * ldb + (uint32_t)bk : (uint32_t)bk * ldb + (uint32_t)b_r0;
    b_soff = (uint32_t)(tb ? 256 / BK : 256 / BN) * ldb;
    bstep = tb ? 1u : ldb;
  };
  auto load = [&](int k0) {
    const uint32_t kb = (uint32_t)k0 * bstep;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const bool ok = a_r0 + i * (256 / BK) < M && a_kt + k0 < segK;
      ra[i] = load_as_f32<float>(ar, ok ? a_off0 + (uint32_t)i * a_soff + (uint32_t)k0 : kOob);
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const bool ok = b_r0 + i * b_sr < N && b_kt + i * b_sk + k0 < segK;
      rb[i] = load_as_f32<TB>(br, ok ? b_off0 + (uint32_t)i * b_soff + kb : kOob);
    }
  };
  auto store = [&](int buf, uint8_t flags) {
    const bool tb = flags & SEG_TRANS_B;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int e = tid + i * 256;
      As[buf][e % BK][e / BK] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int k, n;
      b_kn(i, tb, k, n);
      Bs[buf][k][n] = rb[i];
    }
  };

  const int s_end = tk.seg_begin + tk.seg_count;
  int s = tk.seg_begin;
  while (s < s_end && segs[s].K <= 0) ++s;
  if (s < s_end) {
    GemmSeg cur = segs[s];
    int k0 = 0, buf = 0;
    setup(cur);
    load(0);
    store(0, cur.flags);
    __syncthreads();
    while (true) {
      int ns = s, nk = k0 + BK;
      GemmSeg nxt = cur;
      if (nk >= cur.K) {
        ++ns;
        nk = 0;
        while (ns < s_end && segs[ns].K <= 0) ++ns;
        if (ns < s_end) nxt = segs[ns];
      }
      const bool more = ns < s_end;
      if (more) {
        if (ns != s) setup(nxt);
        load(nk);
      }
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        float af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = As[buf][kk + (lane >> 4)][wm0 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = Bs[buf][kk + (lane >> 4)][wn0 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
      if (!more) break;
      store(buf ^ 1, nxt.flags);
      __syncthreads();
      buf ^= 1;
      s = ns;
      k0 = nk;
      cur = nxt;
    }
  }
  store_tile<FM, FN, TC>(tk, tl, slots, wm0, wn0, lane, acc);
}

// ---------------------------------------------------------------------------
// Split-BF16 slices are 32 deep.  The fragment-split real kernel and the
// complex kernel stage FP32 slices (As[m][k], Bs[n][k], k contiguous) and
// split fragments in registers.
namespace {
constexpr int BKB = 32;
constexpr int KP = BKB + 4;   // row pitch (floats)

template <int NPIECE>
__device__ inline void split8(const float (&x)[8], bf16x8 (&p)[NPIECE]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float r = x[j];
#pragma unroll
    for (int q = 0; q < NPIECE; ++q) {
      const __bf16 h = (__bf16)r;
      p[q][j] = h;
      r -= (float)h;
    }
  }
}
}  // namespace

// Split-BF16, real data, split at the fragment read: LDS holds the FP32
// slices (As[m][k], Bs[n][k], two thirds of the bytes of three piece planes)
// and each wave splits the fragments it reads.  Used for stage 1 (short K:
// few slices per tile, where a second resident workgroup per CU hides the
// tile's prologue and epilogue) and for the 3-term mode (two pieces: less
// split work to save).
template <int BM, int BN, int NPC, typename TB, typename TC>
__global__ __launch_bounds__(256) void gemm_bf16_frag_kernel(const GemmTile *__restrict__ tiles,
                                                        const GemmTask *__restrict__ tasks,
                                                        const GemmSeg *__restrict__ segs, Slots slots, int ntiles) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int A_PER = BM * BKB / 256, B_PER = BKB * BN / 256;
  constexpr int NPIECE = NPC == 3 ? 2 : 3;
  __shared__ float As[2][BM][KP];
  __shared__ float Bs[2][BN][KP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  if ((int)blockIdx.x >= ntiles) return;
  if (slots.skip && *slots.skip) return;   // the Krylov solve has stopped
  const GemmTile tl = tiles[blockIdx.x];
  if (tl.task < 0) return;
  const GemmTask tk = tasks[tl.task];
  const int M = tk.M, N = tk.N;

  flt4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
  float ra[A_PER], rb[B_PER];

  // B element e: consecutive threads read consecutive addresses in either storage order
  auto b_kn = [&](int i, bool tb, int &k, int &n) {
    const int e = tid + i * 256;
    n = tb ? e / BKB : e % BN;
    k = tb ? e % BKB : e / BN;
  };
  // Per-segment addressing (set up at each segment switch).  Since 256 is a
  // multiple of the slice depth and of BN, element i of this thread sits at
  // an affine position: offset off0 + i * soff (plus the slice's k0 term),
  // k = kt + i * sk and row (column) r0 + i * sr, with the steps uniform.
  __amdgpu_buffer_rsrc_t ar, br;
  uint32_t a_off0 = 0, b_off0 = 0, a_soff = 0, b_soff = 0, bstep = 1;
  int a_kt = 0, b_kt = 0, b_sk = 0, a_r0 = 0, b_r0 = 0, b_sr = 0, segK = 0;
  auto setup = [&](const GemmSeg &sg) {
    ar = make_rsrc(reinterpret_cast<const float *>(slots.p[sg.a_slot]) + sg.a_off);
    br = make_rsrc(reinterpret_cast<const TB *>(slots.p[sg.b_slot]) + sg.b_off);
    const uint32_t lda = (uint32_t)sg.lda, ldb = (uint32_t)sg.ldb;
    const bool tb = sg.flags & SEG_TRANS_B;
    segK = sg.K;
    // A element e = tid + 256 i: (m, k) = (e / BKB, e % BKB)
    a_r0 = tl.m0 + tid / BKB;
    a_kt = tid % BKB;
    a_off0 = (uint32_t)a_r0 * lda + (uint32_t)a_kt;
    a_soff = (uint32_t)(256 / BKB) * lda;
    // B element e: (k, n) = (e / BN, e % BN), or (e % BKB, e / BKB) when transposed
    const int bn = tb ? tid / BKB : tid % BN, bk = tb ? tid % BKB : tid / BN;
    b_r0 = tl.n0 + bn;
    b_kt = bk;
    b_sr = tb ? 256 / BKB : 0;
    b_sk = tb ? 0 : 256 / BN;
    b_off0 = tb ? (uint32_t)b_r0 * ldb + (uint32_t)bk : (uint32_t)bk * ldb + (uint32_t)b_r0;
    b_soff = (uint32_t)(tb ? 256 / BKB : 256 / BN) * ldb;
    bstep = tb ? 1u : ldb;
  };
  auto load = [&](int k0) {
    const uint32_t kb = (uint32_t)k0 * bstep;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const bool ok = a_r0 + i * (256 / BKB) < M && a_kt + k0 < segK;
      ra[i] = load_as_f32<float>(ar, ok ? a_off0 + (uint32_t)i * a_soff + (uint32_t)k0 : kOob);
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const bool ok = b_r0 + i * b_sr < N && b_kt + i * b_sk + k0 < segK;
      rb[i] = load_as_f32<TB>(br, ok ? b_off0 + (uint32_t)i * b_soff + kb : kOob);
    }
  };
  auto store = [&](int buf, uint8_t flags) {
    const bool tb = flags & SEG_TRANS_B;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int e = tid + i * 256;
      As[buf][e / BKB][e % BKB] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int k, n;
      b_kn(i, tb, k, n);
      Bs[buf][n][k] = rb[i];
    }
  };
  auto step = [&](int buf) {
    bf16x8 ap[FM][NPIECE], bp[FN][NPIECE];
    const int k8 = 8 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      float x[8];
      const float *src = &As[buf][wm0 + i * 16 + (lane & 15)][k8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = src[j];
      split8<NPIECE>(x, ap[i]);
    }
#pragma unroll
    for (int jn = 0; jn < FN; ++jn) {
      float x[8];
      const float *src = &Bs[buf][wn0 + jn * 16 + (lane & 15)][k8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = src[j];
      split8<NPIECE>(x, bp[jn]);
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        flt4 c = acc[i][j];
        // small products first
        if constexpr (NPC == 6) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][1], bp[j][1], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][2], bp[j][0], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][2], c, 0, 0, 0);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][1], bp[j][0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][0], c, 0, 0, 0);
        acc[i][j] = c;
      }
  };

  const int s_end = tk.seg_begin + tk.seg_count;
  int s = tk.seg_begin;
  while (s < s_end && segs[s].K <= 0) ++s;
  if (s < s_end) {
    GemmSeg cur = segs[s];
    int k0 = 0, buf = 0;
    setup(cur);
    load(0);
    store(0, cur.flags);
    __syncthreads();
    while (true) {
      int ns = s, nk = k0 + BKB;
      GemmSeg nxt = cur;
      if (nk >= cur.K) {
        ++ns;
        nk = 0;
        while (ns < s_end && segs[ns].K <= 0) ++ns;
        if (ns < s_end) nxt = segs[ns];
      }
      const bool more = ns < s_end;
      if (more) {
        if (ns != s) setup(nxt);
        load(nk);
      }
      step(buf);
      if (!more) break;
      store(buf ^ 1, nxt.flags);
      __syncthreads();
      buf ^= 1;
      s = ns;
      k0 = nk;
      cur = nxt;
    }
  }
  store_tile<FM, FN, TC>(tk, tl, slots, wm0, wn0, lane, acc);
}

// Split-BF16, real data.  Each operand value is split into its bf16 pieces
// once, when its K slice is staged: LDS holds one bf16 plane per piece,
// Ap[buf][piece][m][.] and Bp[buf][piece][n][.] (32-deep slices: a row is
// four 16-byte chunks of 8 consecutive k), and every MFMA fragment (8
// consecutive k of one row) is one 16-byte read of a plane.  The chunks of a
// row are stored XOR-swizzled by ((row / 2) ^ (row / 4)) % 4: a row is 64
// bytes, so any 8 consecutive rows of one fragment read land on distinct
// 16-byte slots of a 128-byte bank window and any 16 on distinct slots of a
// 256-byte one (the earlier (row / 4) % 4 left two-way conflicts in every
// group of 8 lanes: 2.7-4 conflict cycles per LDS instruction counted).  Threads stage K in quads of four
// consecutive values: A quads run along a row (coalesced); B quads run along
// n when B is stored [K x N] and along k when it is stored [N x K].
namespace {
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ inline int swz(int row, int chunk) { return chunk ^ (((row >> 1) ^ (row >> 2)) & 3); }

template <int NPIECE>
__device__ inline void split4(const float (&x)[4], bf16x4 (&p)[NPIECE]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float r = x[j];
#pragma unroll
    for (int q = 0; q < NPIECE; ++q) {
      const __bf16 h = (__bf16)r;
      p[q][j] = h;
      r -= (float)h;
    }
  }
}
}  // namespace

template <int BM, int BN, int NPC, typename TB, typename TC>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const GemmTile *__restrict__ tiles,
                                                        const GemmTask *__restrict__ tasks,
                                                        const GemmSeg *__restrict__ segs, Slots slots, int ntiles) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int NPIECE = NPC == 3 ? 2 : 3;
  constexpr int KQ = BKB / 4;                              // quads per row of a slice
  constexpr int QA = BM * KQ / 256, QB = BN * KQ / 256;    // quads per thread
  static_assert(QA >= 1 && QB >= 1 && BM * KQ % 256 == 0 && BN * KQ % 256 == 0, "tile too small for 256 threads");
  __shared__ bf16x8 Ap[2][NPIECE][BM][BKB / 8];
  __shared__ bf16x8 Bp[2][NPIECE][BN][BKB / 8];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  if ((int)blockIdx.x >= ntiles) return;
  if (slots.skip && *slots.skip) return;   // the Krylov solve has stopped
  const GemmTile tl = tiles[blockIdx.x];
  if (tl.task < 0) return;
  const GemmTask tk = tasks[tl.task];
  const int M = tk.M, N = tk.N;

  flt4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
  float ra[QA][4], rb[QB][4];

  // quad i of this thread: row / column and quad index within the slice
  auto a_quad = [&](int i, int &m, int &kq) {
    const int q = tid + i * 256;
    m = q / KQ;
    kq = q % KQ;
  };
  auto b_quad = [&](int i, bool tb, int &n, int &kq) {
    const int q = tid + i * 256;
    n = tb ? q / KQ : q % BN;
    kq = tb ? q % KQ : q / BN;
  };
  auto load = [&](const GemmSeg &sg, int k0) {
    const __amdgpu_buffer_rsrc_t ar = make_rsrc(reinterpret_cast<const float *>(slots.p[sg.a_slot]) + sg.a_off);
    const __amdgpu_buffer_rsrc_t br = make_rsrc(reinterpret_cast<const TB *>(slots.p[sg.b_slot]) + sg.b_off);
    const uint32_t lda = (uint32_t)sg.lda, ldb = (uint32_t)sg.ldb;
    const bool tb = sg.flags & SEG_TRANS_B;
#pragma unroll
    for (int i = 0; i < QA; ++i) {
      int m, kq;
      a_quad(i, m, kq);
      const uint32_t gm = tl.m0 + m;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t gk = k0 + 4 * kq + j;
        ra[i][j] = load_as_f32<float>(ar, ((int)gm < M && (int)gk < sg.K) ? gm * lda + gk : kOob);
      }
    }
#pragma unroll
    for (int i = 0; i < QB; ++i) {
      int n, kq;
      b_quad(i, tb, n, kq);
      const uint32_t gn = tl.n0 + n;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t gk = k0 + 4 * kq + j;
        rb[i][j] = load_as_f32<TB>(br, ((int)gn < N && (int)gk < sg.K) ? (tb ? gn * ldb + gk : gk * ldb + gn) : kOob);
      }
    }
  };
  auto store = [&](int buf, uint8_t flags) {
    const bool tb = flags & SEG_TRANS_B;
#pragma unroll
    for (int i = 0; i < QA; ++i) {
      int m, kq;
      a_quad(i, m, kq);
      bf16x4 pc[NPIECE];
      split4<NPIECE>(ra[i], pc);
#pragma unroll
      for (int p = 0; p < NPIECE; ++p)
        reinterpret_cast<bf16x4 *>(&Ap[buf][p][m][swz(m, kq >> 1)])[kq & 1] = pc[p];
    }
#pragma unroll
    for (int i = 0; i < QB; ++i) {
      int n, kq;
      b_quad(i, tb, n, kq);
      bf16x4 pc[NPIECE];
      split4<NPIECE>(rb[i], pc);
#pragma unroll
      for (int p = 0; p < NPIECE; ++p)
        reinterpret_cast<bf16x4 *>(&Bp[buf][p][n][swz(n, kq >> 1)])[kq & 1] = pc[p];
    }
  };
  auto step = [&](int buf) {
    bf16x8 ap[FM][NPIECE], bp[FN][NPIECE];
    const int c = lane >> 4;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int row = wm0 + i * 16 + (lane & 15);
#pragma unroll
      for (int p = 0; p < NPIECE; ++p) ap[i][p] = Ap[buf][p][row][swz(row, c)];
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int row = wn0 + j * 16 + (lane & 15);
#pragma unroll
      for (int p = 0; p < NPIECE; ++p) bp[j][p] = Bp[buf][p][row][swz(row, c)];
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        flt4 cc = acc[i][j];
        // small products first
        if constexpr (NPC == 6) {
          cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][1], bp[j][1], cc, 0, 0, 0);
          cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][2], bp[j][0], cc, 0, 0, 0);
          cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][2], cc, 0, 0, 0);
        }
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][1], bp[j][0], cc, 0, 0, 0);
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][1], cc, 0, 0, 0);
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][0], cc, 0, 0, 0);
        acc[i][j] = cc;
      }
  };

  const int s_end = tk.seg_begin + tk.seg_count;
  int s = tk.seg_begin;
  while (s < s_end && segs[s].K <= 0) ++s;
  if (s < s_end) {
    GemmSeg cur = segs[s];
    int k0 = 0, buf = 0;
    load(cur, 0);
    store(0, cur.flags);
    __syncthreads();
    while (true) {
      int ns = s, nk = k0 + BKB;
      GemmSeg nxt = cur;
      if (nk >= cur.K) {
        ++ns;
        nk = 0;
        while (ns < s_end && segs[ns].K <= 0) ++ns;
        if (ns < s_end) nxt = segs[ns];
      }
      const bool more = ns < s_end;
      if (more) load(nxt, nk);
      step(buf);
      if (!more) break;
      store(buf ^ 1, nxt.flags);
      __syncthreads();
      buf ^= 1;
      s = ns;
      k0 = nk;
      cur = nxt;
    }
  }
  store_tile<FM, FN, TC>(tk, tl, slots, wm0, wn0, lane, acc);
}

// ---------------------------------------------------------------------------
// Split-BF16, complex data (stored (re, im)): the LDS slices hold real and
// imaginary planes, and three FP32 accumulator planes collect
// Re A . Re B, Im A . Im B and Re A . Im B + Im A . Re B, each from the same
// split-piece products as the real kernel; Re C is the difference of the
// first two, formed once at the end.
template <int FM, int FN, typename TC>
__device__ inline void store_tile_cplx(const GemmTask &tk, const GemmTile &tl, const Slots &slots, int wm0, int wn0,
                                       int lane, const flt4 (&acc)[3][FM][FN]) {
  int64_t crow[FM][4];
  out_rows<FM>(tk, tl, slots, wm0, lane, crow);
  TC *C = reinterpret_cast<TC *>(slots.p[tk.c_slot]) + 2 * tk.c_off;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = tl.n0 + wn0 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = tl.m0 + wm0 + i * 16 + 4 * (lane >> 4) + r;
        if (row < tk.M && col < tk.N) {
          const int64_t ix = 2 * (crow[i][r] + col);
          C[ix] = (TC)(acc[0][i][j][r] - acc[1][i][j][r]);
          C[ix + 1] = (TC)acc[2][i][j][r];
        }
      }
    }
}

template <typename T>
__device__ inline float2 load_cplx_as_f32(__amdgpu_buffer_rsrc_t r, uint32_t elem) {
  if constexpr (sizeof(T) == 8) {
    const double2 v = load_c128(r, elem == kOob ? kOob : elem * 16u);
    return float2{(float)v.x, (float)v.y};
  } else {
    return load_c64(r, elem == kOob ? kOob : elem * 8u);
  }
}

template <int BM, int BN, int NPC, typename TB, typename TC>
__global__ __launch_bounds__(256) void gemm_bf16_cplx_kernel(const GemmTile *__restrict__ tiles,
                                                             const GemmTask *__restrict__ tasks,
                                                             const GemmSeg *__restrict__ segs, Slots slots,
                                                             int ntiles) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int A_PER = BM * BKB / 256, B_PER = BKB * BN / 256;
  constexpr int NPIECE = NPC == 3 ? 2 : 3;
  __shared__ float As[2][2][BM][KP];
  __shared__ float Bs[2][2][BN][KP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  if ((int)blockIdx.x >= ntiles) return;
  if (slots.skip && *slots.skip) return;   // the Krylov solve has stopped
  const GemmTile tl = tiles[blockIdx.x];
  if (tl.task < 0) return;
  const GemmTask tk = tasks[tl.task];
  const int M = tk.M, N = tk.N;

  flt4 acc[3][FM][FN];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[q][i][j] = flt4{0.f, 0.f, 0.f, 0.f};
  float2 ra[A_PER], rb[B_PER];

  auto b_kn = [&](int i, bool tb, int &k, int &n) {
    const int e = tid + i * 256;
    n = tb ? e / BKB : e % BN;
    k = tb ? e % BKB : e / BN;
  };
  // Per-segment addressing (set up at each segment switch).  Since 256 is a
  // multiple of the slice depth and of BN, element i of this thread sits at
  // an affine position: offset off0 + i * soff (plus the slice's k0 term),
  // k = kt + i * sk and row (column) r0 + i * sr, with the steps uniform.
  __amdgpu_buffer_rsrc_t ar, br;
  uint32_t a_off0 = 0, b_off0 = 0, a_soff = 0, b_soff = 0, bstep = 1;
  int a_kt = 0, b_kt = 0, b_sk = 0, a_r0 = 0, b_r0 = 0, b_sr = 0, segK = 0;
  auto setup = [&](const GemmSeg &sg) {
    ar = make_rsrc(reinterpret_cast<const float *>(slots.p[sg.a_slot]) + 2 * sg.a_off);
    br = make_rsrc(reinterpret_cast<const TB *>(slots.p[sg.b_slot]) + 2 * sg.b_off);
    const uint32_t lda = (uint32_t)sg.lda, ldb = (uint32_t)sg.ldb;
    const bool tb = sg.flags & SEG_TRANS_B;
    segK = sg.K;
    // A element e = tid + 256 i: (m, k) = (e / BKB, e % BKB)
    a_r0 = tl.m0 + tid / BKB;
    a_kt = tid % BKB;
    a_off0 = (uint32_t)a_r0 * lda + (uint32_t)a_kt;
    a_soff = (uint32_t)(256 / BKB) * lda;
    // B element e: (k, n) = (e / BN, e % BN), or (e % BKB, e / BKB) when transposed
    const int bn = tb ? tid / BKB : tid % BN, bk = tb ? tid % BKB : tid / BN;
    b_r0 = tl.n0 + bn;
    b_kt = bk;
    b_sr = tb ? 256 / BKB : 0;
    b_sk = tb ? 0 : 256 / BN;
    b_off0 = tb ? (uint32_t)b_r0 * ldb + (uint32_t)bk : (uint32_t)bk * ldb + (uint32_t)b_r0;
    b_soff = (uint32_t)(tb ? 256 / BKB : 256 / BN) * ldb;
    bstep = tb ? 1u : ldb;
  };
  auto load = [&](int k0) {
    const uint32_t kb = (uint32_t)k0 * bstep;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const bool ok = a_r0 + i * (256 / BKB) < M && a_kt + k0 < segK;
      ra[i] = load_cplx_as_f32<float>(ar, ok ? a_off0 + (uint32_t)i * a_soff + (uint32_t)k0 : kOob);
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const bool ok = b_r0 + i * b_sr < N && b_kt + i * b_sk + k0 < segK;
      rb[i] = load_cplx_as_f32<TB>(br, ok ? b_off0 + (uint32_t)i * b_soff + kb : kOob);
    }
  };
  auto store = [&](int buf, uint8_t flags) {
    const bool tb = flags & SEG_TRANS_B;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int e = tid + i * 256;
      As[buf][0][e / BKB][e % BKB] = ra[i].x;
      As[buf][1][e / BKB][e % BKB] = ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int k, n;
      b_kn(i, tb, k, n);
      Bs[buf][0][n][k] = rb[i].x;
      Bs[buf][1][n][k] = rb[i].y;
    }
  };
  auto prod = [&](const bf16x8 (&a)[NPIECE], const bf16x8 (&b)[NPIECE], flt4 c) {
    if constexpr (NPC == 6) {
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], c, 0, 0, 0);
    }
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], c, 0, 0, 0);
    return c;
  };
  auto step = [&](int buf) {
    bf16x8 ap[2][FM][NPIECE], bp[2][FN][NPIECE];
    const int k8 = 8 * (lane >> 4);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        float x[8];
        const float *src = &As[buf][p][wm0 + i * 16 + (lane & 15)][k8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = src[j];
        split8<NPIECE>(x, ap[p][i]);
      }
#pragma unroll
      for (int jn = 0; jn < FN; ++jn) {
        float x[8];
        const float *src = &Bs[buf][p][wn0 + jn * 16 + (lane & 15)][k8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = src[j];
        split8<NPIECE>(x, bp[p][jn]);
      }
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        acc[0][i][j] = prod(ap[0][i], bp[0][j], acc[0][i][j]);
        acc[1][i][j] = prod(ap[1][i], bp[1][j], acc[1][i][j]);
        acc[2][i][j] = prod(ap[1][i], bp[0][j], prod(ap[0][i], bp[1][j], acc[2][i][j]));
      }
  };

  const int s_end = tk.seg_begin + tk.seg_count;
  int s = tk.seg_begin;
  while (s < s_end && segs[s].K <= 0) ++s;
  if (s < s_end) {
    GemmSeg cur = segs[s];
    int k0 = 0, buf = 0;
    setup(cur);
    load(0);
    store(0, cur.flags);
    __syncthreads();
    while (true) {
      int ns = s, nk = k0 + BKB;
      GemmSeg nxt = cur;
      if (nk >= cur.K) {
        ++ns;
        nk = 0;
        while (ns < s_end && segs[ns].K <= 0) ++ns;
        if (ns < s_end) nxt = segs[ns];
      }
      const bool more = ns < s_end;
      if (more) {
        if (ns != s) setup(nxt);
        load(nk);
      }
      step(buf);
      if (!more) break;
      store(buf ^ 1, nxt.flags);
      __syncthreads();
      buf ^= 1;
      s = ns;
      k0 = nk;
      cur = nxt;
    }
  }
  store_tile_cplx<FM, FN, TC>(tk, tl, slots, wm0, wn0, lane, acc);
}

// ---------------------------------------------------------------------------
// Split-BF16 on pre-split operands.  The A operand (and B when BPLANAR) is
// held in memory as bf16 piece planes (h, m, l of each FP32 value, plane q at
// ps[slot] elements after plane q - 1), laid out like the FP32 operand with a
// leading dimension that is a multiple of 8 and zero padding past K in every
// row.  A K slice is then staged by 16-byte loads of 8 consecutive k values
// per plane, straight into the LDS piece planes: no split work.  A B operand
// that is not planar (stage 1's B, the FP64 Krylov vector) is split when its
// slice is staged, as in gemm_bf16_kernel.  The MFMA step is the same as
// there.  Chunks of 8 k values that start at or past K read nothing (kOob);
// a chunk that straddles K reads the row's zero padding.
// LDS slices per workgroup in the pre-split kernel: one.  A ping-pong pair
// saves a barrier per slice but doubles the LDS footprint (three planes per
// operand), which held the 128x64 and 64x64 six-term tiles to 2-3 workgroups
// per CU.  With one slice the next slice's loads still overlap this slice's
// MFMAs (they land in registers first); the six-term stage 1 measured 18 %
// faster, stage 3 the same.
constexpr int kSliceBufs = 1;

template <int BM, int BN, int NPC, typename TB, typename TC, bool BPLANAR>
__global__ __launch_bounds__(256) void gemm_bf16p_kernel(const GemmTile *__restrict__ tiles,
                                                         const GemmTask *__restrict__ tasks,
                                                         const GemmSeg *__restrict__ segs, Slots slots, int ntiles) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int NPIECE = NPC == 3 ? 2 : 3;
  constexpr int KC8 = BKB / 8;                                   // 8-value chunks per row of a slice
  constexpr int CA = (BM * KC8 + 255) / 256, CB = (BN * KC8 + 255) / 256;   // chunks per thread (planar)
  constexpr int KQ = BKB / 4;
  constexpr int QB = BN * KQ / 256;                              // quads per thread (split B)
  static_assert(BPLANAR || (QB >= 1 && BN * KQ % 256 == 0), "tile too small for 256 threads");
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  constexpr int NB = kSliceBufs;   // 1: one LDS slice; 2: a ping-pong pair
  __shared__ bf16x8 Ap[NB][NPIECE][BM][BKB / 8];
  __shared__ bf16x8 Bp[NB][NPIECE][BN][BKB / 8];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  if ((int)blockIdx.x >= ntiles) return;
  if (slots.skip && *slots.skip) return;   // the Krylov solve has stopped
  const GemmTile tl = tiles[blockIdx.x];
  if (tl.task < 0) return;
  const GemmTask tk = tasks[tl.task];
  const int M = tk.M, N = tk.N;

  flt4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
  u4 ra[NPIECE][CA];
  u4 rbp[BPLANAR ? NPIECE : 1][BPLANAR ? CB : 1];
  float rb[BPLANAR ? 1 : QB][4];

  auto load = [&](const GemmSeg &sg, int k0) {
    // A: planar, [M x K] rows with leading dimension lda
    {
      const __bf16 *a0 = reinterpret_cast<const __bf16 *>(slots.p[sg.a_slot]) + sg.a_off;
      const int64_t aps = slots.ps[sg.a_slot];
      const uint32_t lda = (uint32_t)sg.lda;
#pragma unroll
      for (int p = 0; p < NPIECE; ++p) {
        const __amdgpu_buffer_rsrc_t ar = make_rsrc(a0 + p * aps);
#pragma unroll
        for (int i = 0; i < CA; ++i) {
          const int q = tid + i * 256;
          const int m = q / KC8, c = q % KC8;
          const uint32_t gm = tl.m0 + m, gk = k0 + 8 * c;
          const bool ok = q < BM * KC8 && (int)gm < M && (int)gk < sg.K;
          ra[p][i] = __builtin_amdgcn_raw_buffer_load_b128(ar, ok ? (gm * lda + gk) * 2u : kOob, 0, 0);
        }
      }
    }
    if constexpr (BPLANAR) {
      // B: planar, stored [N x K] (transposed operand) with leading dimension ldb
      const __bf16 *b0 = reinterpret_cast<const __bf16 *>(slots.p[sg.b_slot]) + sg.b_off;
      const int64_t bps = slots.ps[sg.b_slot];
      const uint32_t ldb = (uint32_t)sg.ldb;
#pragma unroll
      for (int p = 0; p < NPIECE; ++p) {
        const __amdgpu_buffer_rsrc_t br = make_rsrc(b0 + p * bps);
#pragma unroll
        for (int i = 0; i < CB; ++i) {
          const int q = tid + i * 256;
          const int n = q / KC8, c = q % KC8;
          const uint32_t gn = tl.n0 + n, gk = k0 + 8 * c;
          const bool ok = q < BN * KC8 && (int)gn < N && (int)gk < sg.K;
          rbp[p][i] = __builtin_amdgcn_raw_buffer_load_b128(br, ok ? (gn * ldb + gk) * 2u : kOob, 0, 0);
        }
      }
    } else {
      // B: TB values stored [K x N] (or [N x K] when transposed), split at staging
      const __amdgpu_buffer_rsrc_t br = make_rsrc(reinterpret_cast<const TB *>(slots.p[sg.b_slot]) + sg.b_off);
      const uint32_t ldb = (uint32_t)sg.ldb;
      const bool tb = sg.flags & SEG_TRANS_B;
#pragma unroll
      for (int i = 0; i < QB; ++i) {
        const int q = tid + i * 256;
        const int n = tb ? q / KQ : q % BN, kq = tb ? q % KQ : q / BN;
        const uint32_t gn = tl.n0 + n;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t gk = k0 + 4 * kq + j;
          rb[i][j] = load_as_f32<TB>(br, ((int)gn < N && (int)gk < sg.K) ? (tb ? gn * ldb + gk : gk * ldb + gn) : kOob);
        }
      }
    }
  };
  auto store = [&](int buf, uint8_t flags) {
#pragma unroll
    for (int p = 0; p < NPIECE; ++p)
#pragma unroll
      for (int i = 0; i < CA; ++i) {
        const int q = tid + i * 256;
        if (q < BM * KC8) {
          const int m = q / KC8, c = q % KC8;
          Ap[buf][p][m][swz(m, c)] = __builtin_bit_cast(bf16x8, ra[p][i]);
        }
      }
    if constexpr (BPLANAR) {
#pragma unroll
      for (int p = 0; p < NPIECE; ++p)
#pragma unroll
        for (int i = 0; i < CB; ++i) {
          const int q = tid + i * 256;
          if (q < BN * KC8) {
            const int n = q / KC8, c = q % KC8;
            Bp[buf][p][n][swz(n, c)] = __builtin_bit_cast(bf16x8, rbp[p][i]);
          }
        }
    } else {
      const bool tb = flags & SEG_TRANS_B;
#pragma unroll
      for (int i = 0; i < QB; ++i) {
        const int q = tid + i * 256;
        const int n = tb ? q / KQ : q % BN, kq = tb ? q % KQ : q / BN;
        bf16x4 pc[NPIECE];
        split4<NPIECE>(rb[i], pc);
#pragma unroll
        for (int p = 0; p < NPIECE; ++p)
          reinterpret_cast<bf16x4 *>(&Bp[buf][p][n][swz(n, kq >> 1)])[kq & 1] = pc[p];
      }
    }
  };
  auto step = [&](int buf) {
    bf16x8 ap[FM][NPIECE], bp[FN][NPIECE];
    const int c = lane >> 4;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int row = wm0 + i * 16 + (lane & 15);
#pragma unroll
      for (int p = 0; p < NPIECE; ++p) ap[i][p] = Ap[buf][p][row][swz(row, c)];
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int row = wn0 + j * 16 + (lane & 15);
#pragma unroll
      for (int p = 0; p < NPIECE; ++p) bp[j][p] = Bp[buf][p][row][swz(row, c)];
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        flt4 cc = acc[i][j];
        if constexpr (NPC == 6) {
          cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][1], bp[j][1], cc, 0, 0, 0);
          cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][2], bp[j][0], cc, 0, 0, 0);
          cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][2], cc, 0, 0, 0);
        }
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][1], bp[j][0], cc, 0, 0, 0);
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][1], cc, 0, 0, 0);
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i][0], bp[j][0], cc, 0, 0, 0);
        acc[i][j] = cc;
      }
  };

  const int s_end = tk.seg_begin + tk.seg_count;
  int s = tk.seg_begin;
  while (s < s_end && segs[s].K <= 0) ++s;
  if (s < s_end) {
    GemmSeg cur = segs[s];
    int k0 = 0, buf = 0;
    load(cur, 0);
    store(0, cur.flags);
    __syncthreads();
    while (true) {
      int ns = s, nk = k0 + BKB;
      GemmSeg nxt = cur;
      if (nk >= cur.K) {
        ++ns;
        nk = 0;
        while (ns < s_end && segs[ns].K <= 0) ++ns;
        if (ns < s_end) nxt = segs[ns];
      }
      const bool more = ns < s_end;
      if (more) load(nxt, nk);
      step(buf);
      if (!more) break;
      if constexpr (NB == 1) __syncthreads();   // every wave has read the slice
      store(NB == 1 ? 0 : buf ^ 1, nxt.flags);
      __syncthreads();
      if constexpr (NB == 2) buf ^= 1;
      s = ns;
      k0 = nk;
      cur = nxt;
    }
  }
  store_tile<FM, FN, TC>(tk, tl, slots, wm0, wn0, lane, acc);
}

// FP64 -> NPIECE bf16 planes of the FP32 value (h, m, l), element for element:
// the pre-split constant operands, made once per plan
template <int NPIECE>
__global__ __launch_bounds__(256) void split_planes_kernel(const double *__restrict__ src, __bf16 *__restrict__ dst,
                                                           int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    float r = (float)src[i];
#pragma unroll
    for (int q = 0; q < NPIECE; ++q) {
      const __bf16 h = (__bf16)r;
      dst[q * n + i] = h;
      r -= (float)h;
    }
  }
}

void launch_split_planes(const double *src, void *dst, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(split_planes_kernel<3>, dim3((unsigned)blocks), dim3(256), 0, st, src,
                     static_cast<__bf16 *>(dst), n);
  QLG_HIP(hipGetLastError());
}

// stage 1 (A planar, B = the FP64 vector split at staging, FP32 out) or
// stage 3 (A and B planar, FP64 out through the row table)
void launch_grouped_gemm_planar(int mode, int bm, int bn, bool stage3, const GemmTile *tiles, const GemmTask *tasks,
                                const GemmSeg *segs, const Slots &slots, int ntiles, hipStream_t st) {
  if (ntiles <= 0) return;
  dim3 grid(ntiles), block(256);
#define QLG_BP(M_, N_, P_)                                                                                      \
  if (bm == M_ && bn == N_) {                                                                                   \
    if (stage3)                                                                                                 \
      hipLaunchKernelGGL((gemm_bf16p_kernel<M_, N_, P_, float, double, true>), grid, block, 0, st, tiles, tasks, \
                         segs, slots, ntiles);                                                                  \
    else                                                                                                        \
      hipLaunchKernelGGL((gemm_bf16p_kernel<M_, N_, P_, double, float, false>), grid, block, 0, st, tiles,      \
                         tasks, segs, slots, ntiles);                                                           \
    QLG_HIP(hipGetLastError());                                                                                 \
    return;                                                                                                     \
  }
  if (mode == QLG_PREC_BF16X3) {
    QLG_BP(32, 32, 3) QLG_BP(64, 32, 3) QLG_BP(32, 64, 3) QLG_BP(64, 64, 3) QLG_BP(128, 64, 3)
    QLG_BP(64, 128, 3) QLG_BP(128, 128, 3)
  } else if (mode == QLG_PREC_BF16X6) {
    QLG_BP(32, 32, 6) QLG_BP(64, 32, 6) QLG_BP(32, 64, 6) QLG_BP(64, 64, 6) QLG_BP(128, 64, 6)
    QLG_BP(64, 128, 6) QLG_BP(128, 128, 6)
  }
#undef QLG_BP
  throw Error(QLG_ERR_UNSUPPORTED, "unsupported split-BF16 tile shape or mode");
}

// ---------------------------------------------------------------------------
// launchers.  b_f64: B operand FP64 (else FP32); c_f64: output FP64 (else FP32).
template <typename TB, typename TC>
static void launch_mixed_cplx(int mode, int bm, int bn, const GemmTile *tiles, const GemmTask *tasks,
                              const GemmSeg *segs, const Slots &slots, int ntiles, hipStream_t st) {
  dim3 grid(ntiles), block(256);
#define QLG_BFC(M_, N_, P_)                                                                                   \
  if (bm == M_ && bn == N_) {                                                                                 \
    hipLaunchKernelGGL((gemm_bf16_cplx_kernel<M_, N_, P_, TB, TC>), grid, block, 0, st, tiles, tasks, segs,   \
                       slots, ntiles);                                                                        \
    return;                                                                                                   \
  }
  if (mode == QLG_PREC_BF16X3) {
    QLG_BFC(32, 32, 3) QLG_BFC(64, 32, 3) QLG_BFC(32, 64, 3) QLG_BFC(64, 64, 3)
  } else if (mode == QLG_PREC_BF16X6) {
    QLG_BFC(32, 32, 6) QLG_BFC(64, 32, 6) QLG_BFC(32, 64, 6) QLG_BFC(64, 64, 6)
  }
#undef QLG_BFC
  throw Error(QLG_ERR_UNSUPPORTED, "unsupported complex reduced-precision tile shape or mode");
}

template <typename TB, typename TC>
static void launch_mixed_typed(int mode, int bm, int bn, const GemmTile *tiles, const GemmTask *tasks,
                               const GemmSeg *segs, const Slots &slots, int ntiles, hipStream_t st) {
  dim3 grid(ntiles), block(256);
#define QLG_F32(M_, N_)                                                                                       \
  if (bm == M_ && bn == N_) {                                                                                 \
    hipLaunchKernelGGL((gemm_f32_kernel<M_, N_, 16, TB, TC>), grid, block, 0, st, tiles, tasks, segs, slots,  \
                       ntiles);                                                                               \
    return;                                                                                                   \
  }
// Split once at staging only for stage 3 (FP32 operands, long K) in the
// 6-term mode, where the split work is largest; stage 1 (B = the FP64 Krylov
// vector, short K) and the 3-term mode split at the fragment read (measured
// faster at D=2000 for both the synthetic and a converged Hubbard state).
#define QLG_BF(M_, N_, P_)                                                                                    \
  if (bm == M_ && bn == N_) {                                                                                 \
    if constexpr (sizeof(TB) == 8 || P_ == 3)                                                                 \
      hipLaunchKernelGGL((gemm_bf16_frag_kernel<M_, N_, P_, TB, TC>), grid, block, 0, st, tiles, tasks, segs, \
                         slots, ntiles);                                                                      \
    else                                                                                                      \
      hipLaunchKernelGGL((gemm_bf16_kernel<M_, N_, P_, TB, TC>), grid, block, 0, st, tiles, tasks, segs,      \
                         slots, ntiles);                                                                      \
    return;                                                                                                   \
  }
  if (mode == QLG_PREC_F32) {
    QLG_F32(32, 32) QLG_F32(64, 32) QLG_F32(32, 64) QLG_F32(64, 64) QLG_F32(128, 64) QLG_F32(64, 128)
    QLG_F32(128, 128)
  } else if (mode == QLG_PREC_BF16X3) {
    QLG_BF(32, 32, 3) QLG_BF(64, 32, 3) QLG_BF(32, 64, 3) QLG_BF(64, 64, 3) QLG_BF(128, 64, 3)
    QLG_BF(64, 128, 3) QLG_BF(128, 128, 3)
  } else if (mode == QLG_PREC_BF16X6) {
    QLG_BF(32, 32, 6) QLG_BF(64, 32, 6) QLG_BF(32, 64, 6) QLG_BF(64, 64, 6) QLG_BF(128, 64, 6)
    QLG_BF(64, 128, 6) QLG_BF(128, 128, 6)
  }
#undef QLG_F32
#undef QLG_BF
  throw Error(QLG_ERR_UNSUPPORTED, "unsupported reduced-precision tile shape or mode");
}

__global__ __launch_bounds__(256) void f64_to_f32_kernel(const double *__restrict__ src, float *__restrict__ dst,
                                                         int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    dst[i] = (float)src[i];
}

void launch_f64_to_f32(const double *src, float *dst, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
  QLG_HIP(hipGetLastError());
}

void launch_grouped_gemm_mixed(int mode, int bm, int bn, bool b_f64, bool c_f64, bool cplx, const GemmTile *tiles,
                               const GemmTask *tasks, const GemmSeg *segs, const Slots &slots, int ntiles,
                               hipStream_t st) {
  if (ntiles <= 0) return;
  if (cplx) {
    if (b_f64 && !c_f64)
      launch_mixed_cplx<double, float>(mode, bm, bn, tiles, tasks, segs, slots, ntiles, st);
    else if (!b_f64 && c_f64)
      launch_mixed_cplx<float, double>(mode, bm, bn, tiles, tasks, segs, slots, ntiles, st);
    else
      throw Error(QLG_ERR_UNSUPPORTED, "reduced-precision GEMM operand combination");
    QLG_HIP(hipGetLastError());
    return;
  }
  if (b_f64 && !c_f64)
    launch_mixed_typed<double, float>(mode, bm, bn, tiles, tasks, segs, slots, ntiles, st);
  else if (!b_f64 && c_f64)
    launch_mixed_typed<float, double>(mode, bm, bn, tiles, tasks, segs, slots, ntiles, st);
  else
    throw Error(QLG_ERR_UNSUPPORTED, "reduced-precision GEMM operand combination");
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
