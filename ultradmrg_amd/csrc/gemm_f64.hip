// Persistent-free grouped FP64 GEMM on CDNA4 MFMA (v_mfma_f64_16x16x4_f64).
//
// One workgroup (4 wavefronts, 2x2) computes one BM x BN output tile of one
// task; it walks the task's K-segments in order, staging BK-deep slices of A
// and B through LDS with a register prefetch of the next slice.  Each wave
// owns a (BM/2) x (BN/2) sub-tile made of 16x16 MFMA blocks.
//
// MFMA f64 16x16x4 operand layout (one double per lane):
//   A: lane l holds A[row = l % 16][k = l / 16]
//   B: lane l holds B[k = l / 16][col = l % 16]
//   D: lane l, register r holds D[row = l / 16 + 4 * r][col = l % 16]
//      (measured on gfx950 with an in-bounds probe of the builtin)
#include <hip/hip_runtime.h>

#include "common.h"
#include "gemm.h"

namespace qlg {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int BK = 16;
constexpr int PAD = 1;

template <int BM, int BN, bool CPLX>
__global__ __launch_bounds__(256) void grouped_gemm_f64_kernel(const GemmTile *__restrict__ tiles,
                                                               const GemmTask *__restrict__ tasks,
                                                               const GemmSeg *__restrict__ segs,
                                                               Slots slots, int ntiles) {
  constexpr int WM = BM / 2, WN = BN / 2;       // wave sub-tile
  constexpr int FM = WM / 16, FN = WN / 16;     // MFMA blocks per wave
  constexpr int A_PER = BM * BK / 256;          // A elements loaded per thread
  constexpr int B_PER = BK * BN / 256;
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small for 256 threads");
  constexpr int NP = CPLX ? 2 : 1;              // real / imaginary planes
  __shared__ double As[NP][BK][BM + PAD];
  __shared__ double Bs[NP][BK][BN + PAD];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const int tile_id = blockIdx.x;
  if (tile_id >= ntiles) return;
  const GemmTile tl = tiles[tile_id];
  if (tl.task < 0) return;  // padding slot of the XCD-balanced launch order
  const GemmTask tk = tasks[tl.task];
  const int M = tk.M, N = tk.N;

  dbl4 acc[NP][FM][FN];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[p][i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

  double ra[NP][A_PER], rb[NP][B_PER];

  // load the BK slice starting at k0 of segment s into registers; element e of
  // a thread maps to (row, k) so that consecutive lanes read consecutive
  // addresses of the operand's storage order
  auto load = [&](const GemmSeg &sg, int k0) {
    const double *A = slots.p[sg.a_slot] + sg.a_off * NP;
    const double *B = slots.p[sg.b_slot] + sg.b_off * NP;
    const double sa = (sg.flags & SEG_CONJ_A) ? -1.0 : 1.0;
    const double sb = (sg.flags & SEG_CONJ_B) ? -1.0 : 1.0;
    const bool ta = sg.flags & SEG_TRANS_A, tb = sg.flags & SEG_TRANS_B;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int e = tid + i * 256;
      const int m = ta ? e % BM : e / BK, k = ta ? e / BM : e % BK;
      const int gm = tl.m0 + m, gk = k0 + k;
      const bool ok = gm < M && gk < sg.K;
      const int64_t ix = ta ? (int64_t)gk * sg.lda + gm : (int64_t)gm * sg.lda + gk;
      if constexpr (CPLX) {
        const double2 v = ok ? reinterpret_cast<const double2 *>(A)[ix] : double2{0.0, 0.0};
        ra[0][i] = v.x;
        ra[1][i] = sa * v.y;
      } else {
        ra[0][i] = ok ? A[ix] : 0.0;
      }
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int e = tid + i * 256;
      const int n = tb ? e / BK : e % BN, k = tb ? e % BK : e / BN;
      const int gn = tl.n0 + n, gk = k0 + k;
      const bool ok = gn < N && gk < sg.K;
      const int64_t ix = tb ? (int64_t)gn * sg.ldb + gk : (int64_t)gk * sg.ldb + gn;
      if constexpr (CPLX) {
        const double2 v = ok ? reinterpret_cast<const double2 *>(B)[ix] : double2{0.0, 0.0};
        rb[0][i] = v.x;
        rb[1][i] = sb * v.y;
      } else {
        rb[0][i] = ok ? B[ix] : 0.0;
      }
    }
  };
  auto store = [&](uint8_t flags) {
    const bool ta = flags & SEG_TRANS_A, tb = flags & SEG_TRANS_B;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int e = tid + i * 256;
      const int m = ta ? e % BM : e / BK, k = ta ? e / BM : e % BK;
#pragma unroll
      for (int p = 0; p < NP; ++p) As[p][k][m] = ra[p][i];
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int e = tid + i * 256;
      const int n = tb ? e / BK : e % BN, k = tb ? e % BK : e / BN;
#pragma unroll
      for (int p = 0; p < NP; ++p) Bs[p][k][n] = rb[p][i];
    }
  };

  const int s_end = tk.seg_begin + tk.seg_count;
  int s = tk.seg_begin, k0 = 0;
  // skip empty segments
  while (s < s_end && segs[s].K <= 0) ++s;
  if (s < s_end) {
    GemmSeg cur = segs[s];
    load(cur, 0);
    while (true) {
      const uint8_t cur_flags = cur.flags;
      __syncthreads();
      store(cur_flags);
      __syncthreads();
      // advance iterator and prefetch the next slice
      int ns = s, nk = k0 + BK;
      GemmSeg nxt = cur;
      if (nk >= cur.K) {
        ++ns;
        nk = 0;
        while (ns < s_end && segs[ns].K <= 0) ++ns;
        if (ns < s_end) nxt = segs[ns];
      }
      const bool more = ns < s_end;
      if (more) load(nxt, nk);
      const int krem = cur.K - k0;  // valid k of the staged slice; the rest is zero fill
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        if (kk >= krem) break;
        double af[NP][FM], bf[NP][FN];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
          for (int i = 0; i < FM; ++i) af[p][i] = As[p][kk + (lane >> 4)][wm0 + i * 16 + (lane & 15)];
#pragma unroll
          for (int j = 0; j < FN; ++j) bf[p][j] = Bs[p][kk + (lane >> 4)][wn0 + j * 16 + (lane & 15)];
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[0][i], bf[0][j], acc[0][i][j], 0, 0, 0);
            if constexpr (CPLX) {
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-af[1][i], bf[1][j], acc[0][i][j], 0, 0, 0);
              acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[0][i], bf[1][j], acc[1][i][j], 0, 0, 0);
              acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[1][i], bf[0][j], acc[1][i][j], 0, 0, 0);
            }
          }
      }
      if (!more) break;
      s = ns;
      k0 = nk;
      cur = nxt;
    }
  }

  double *C = slots.p[tk.c_slot] + tk.c_off * NP;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = tl.n0 + wn0 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = tl.m0 + wm0 + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col < N) {
          const int64_t ix = (int64_t)row * tk.ldc + col;
          if constexpr (CPLX)
            reinterpret_cast<double2 *>(C)[ix] = double2{acc[0][i][j][r], acc[1][i][j][r]};
          else
            C[ix] = acc[0][i][j][r];
        }
      }
    }
}

template <bool CPLX>
static void launch_shape(int bm, int bn, const GemmTile *tiles, const GemmTask *tasks, const GemmSeg *segs,
                         const Slots &slots, int ntiles, hipStream_t st) {
  dim3 grid(ntiles), block(256);
#define QLG_SHAPE(M_, N_)                                                                                    \
  if (bm == M_ && bn == N_) {                                                                                \
    hipLaunchKernelGGL((grouped_gemm_f64_kernel<M_, N_, CPLX>), grid, block, 0, st, tiles, tasks, segs, slots, \
                       ntiles);                                                                              \
    return;                                                                                                  \
  }
  QLG_SHAPE(32, 32)
  QLG_SHAPE(64, 32)
  QLG_SHAPE(32, 64)
  QLG_SHAPE(64, 64)
  QLG_SHAPE(128, 64)
  QLG_SHAPE(64, 128)
  QLG_SHAPE(128, 128)
#undef QLG_SHAPE
  throw Error(QLG_ERR_UNSUPPORTED, "unsupported FP64 tile shape");
}

void launch_grouped_gemm_f64(int bm, int bn, bool cplx, const GemmTile *tiles, const GemmTask *tasks,
                             const GemmSeg *segs, const Slots &slots, int ntiles, hipStream_t st) {
  if (ntiles <= 0) return;
  if (cplx)
    launch_shape<true>(bm, bn, tiles, tasks, segs, slots, ntiles, st);
  else
    launch_shape<false>(bm, bn, tiles, tasks, segs, slots, ntiles, st);
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
