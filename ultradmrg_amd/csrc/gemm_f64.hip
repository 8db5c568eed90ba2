// Grouped FP64 GEMM on CDNA4 MFMA (v_mfma_f64_16x16x4_f64).
//
// One workgroup (4 wavefronts, 2x2) computes one BM x BN output tile of one
// task and walks the task's K-segments in order.  BK-deep slices of A and B
// are staged through two LDS buffers (ping-pong): the raw buffer loads of
// slice i+1 are issued before the MFMAs of slice i and land in registers,
// which are written to the other LDS buffer afterwards, so each K step costs
// one barrier and the load latency hides behind the MFMA work.  Operands are
// read with raw buffer loads whose out-of-range elements come back as zero
// (tile_io.h), so the loads carry no branch and no select.  Slices run as
// straight-line code, which lets the LDS fragment reads of one 4-deep
// sub-step overlap the MFMAs of the previous one (a segment's last slice is
// zero-filled past K and runs the same code).  Each wave owns a
// (BM/2) x (BN/2) sub-tile of 16x16 MFMA blocks; thin tiles (BN = 16, for
// tasks with few columns) stack the four waves along M.
//
// MFMA f64 16x16x4 operand layout (one double per lane):
//   A: lane l holds A[row = l % 16][k = l / 16]
//   B: lane l holds B[k = l / 16][col = l % 16]
//   D: lane l, register r holds D[row = l / 16 + 4 * r][col = l % 16]
//      (measured on gfx950 with an in-bounds probe of the builtin)
#include <hip/hip_runtime.h>

#include "common.h"
#include "gemm.h"
#include "tile_io.h"

namespace qlg {

typedef double dbl4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int PAD = 1;
}

template <int BM, int BN, int BK, bool CPLX>
__global__ __launch_bounds__(256) void grouped_gemm_f64_kernel(const GemmTile *__restrict__ tiles,
                                                               const GemmTask *__restrict__ tasks,
                                                               const GemmSeg *__restrict__ segs,
                                                               Slots slots, int ntiles) {
  // four waves: 2 x 2, or 4 x 1 for thin tiles (BN = 16)
  constexpr int WAVES_N = BN >= 32 ? 2 : 1, WAVES_M = 4 / WAVES_N;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;   // wave sub-tile
  constexpr int FM = WM / 16, FN = WN / 16;     // MFMA blocks per wave
  constexpr int A_PER = BM * BK / 256;          // A elements loaded per thread per slice
  constexpr int B_PER = BK * BN / 256;
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small for 256 threads");
  static_assert(WM >= 16 && WN >= 16, "wave sub-tile below one MFMA block");
  constexpr int NP = CPLX ? 2 : 1;              // real / imaginary planes
  constexpr uint32_t EB = CPLX ? 16u : 8u;      // element bytes
  __shared__ double As[2][NP][BK][BM + PAD];
  __shared__ double Bs[2][NP][BK][BN + PAD];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm0 = (wave / WAVES_N) * WM, wn0 = (wave % WAVES_N) * WN;
  if ((int)blockIdx.x >= ntiles) return;
  if (slots.skip && *slots.skip) return;   // the Krylov solve has stopped
  const GemmTile tl = tiles[blockIdx.x];
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

  // element e of a slice: (m, k) of A and (k, n) of B, chosen so that
  // consecutive threads read consecutive addresses in either storage order.
  // The transposition flags are uniform, so the choice is a select on the
  // indices and every slice issues one straight sequence of loads.
  auto a_mk = [&](int i, bool ta, int &m, int &k) {
    const int e = tid + i * 256;
    m = ta ? e % BM : e / BK;
    k = ta ? e / BM : e % BK;
  };
  auto b_kn = [&](int i, bool tb, int &k, int &n) {
    const int e = tid + i * 256;
    n = tb ? e / BK : e % BN;
    k = tb ? e % BK : e / BN;
  };
  // Per-segment state of the slices being loaded: buffer descriptors, and
  // each element's offset at k0 = 0, its k within the slice and whether its
  // row / column lies inside the task.  A slice then costs one add, one
  // compare and one select per element.
  __amdgpu_buffer_rsrc_t ar, br;
  uint32_t astep = 1, bstep = 1;
  int segK = 0;
  uint32_t aoff[A_PER], boff[B_PER];
  int akk[A_PER], bkk[B_PER];
  bool ain[A_PER], bin[B_PER];
  auto setup = [&](const GemmSeg &sg) {
    ar = make_rsrc(slots.p[sg.a_slot] + sg.a_off * NP);
    br = make_rsrc(slots.p[sg.b_slot] + sg.b_off * NP);
    const uint32_t lda = (uint32_t)sg.lda, ldb = (uint32_t)sg.ldb;
    const bool ta = sg.flags & SEG_TRANS_A, tb = sg.flags & SEG_TRANS_B;
    segK = sg.K;
    astep = ta ? lda : 1u;
    bstep = tb ? 1u : ldb;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int m, k;
      a_mk(i, ta, m, k);
      const uint32_t gm = tl.m0 + m;
      aoff[i] = ta ? (uint32_t)k * lda + gm : gm * lda + (uint32_t)k;
      akk[i] = k;
      ain[i] = (int)gm < M;
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int k, n;
      b_kn(i, tb, k, n);
      const uint32_t gn = tl.n0 + n;
      boff[i] = tb ? gn * ldb + (uint32_t)k : (uint32_t)k * ldb + gn;
      bkk[i] = k;
      bin[i] = (int)gn < N;
    }
  };
  auto load = [&](int k0) {
    const uint32_t ka = (uint32_t)k0 * astep, kb = (uint32_t)k0 * bstep;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const uint32_t off = (ain[i] && akk[i] + k0 < segK) ? (aoff[i] + ka) * EB : kOob;
      if constexpr (CPLX) {
        const double2 v = load_c128(ar, off);
        ra[0][i] = v.x;
        ra[1][i] = v.y;
      } else {
        ra[0][i] = load_f64(ar, off);
      }
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const uint32_t off = (bin[i] && bkk[i] + k0 < segK) ? (boff[i] + kb) * EB : kOob;
      if constexpr (CPLX) {
        const double2 v = load_c128(br, off);
        rb[0][i] = v.x;
        rb[1][i] = v.y;
      } else {
        rb[0][i] = load_f64(br, off);
      }
    }
  };
  auto store = [&](int buf, uint8_t flags) {
    const bool ta = flags & SEG_TRANS_A, tb = flags & SEG_TRANS_B;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int m, k;
      a_mk(i, ta, m, k);
      As[buf][0][k][m] = ra[0][i];
      if constexpr (CPLX) As[buf][1][k][m] = (flags & SEG_CONJ_A) ? -ra[1][i] : ra[1][i];
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int k, n;
      b_kn(i, tb, k, n);
      Bs[buf][0][k][n] = rb[0][i];
      if constexpr (CPLX) Bs[buf][1][k][n] = (flags & SEG_CONJ_B) ? -rb[1][i] : rb[1][i];
    }
  };
  // one 4-deep sub-step of the staged slice
  auto step = [&](int buf, int kk) {
    double af[NP][FM], bf[NP][FN];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int i = 0; i < FM; ++i) af[p][i] = As[buf][p][kk + (lane >> 4)][wm0 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[p][j] = Bs[buf][p][kk + (lane >> 4)][wn0 + j * 16 + (lane & 15)];
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
      // iterator of the next slice; its loads stay in flight during this slice's MFMAs
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
      // the staged slice is zero-filled past the segment's K, so a partial
      // slice runs the same straight-line code as a full one
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) step(buf, kk);
      if (!more) break;
      store(buf ^ 1, nxt.flags);
      __syncthreads();
      buf ^= 1;
      s = ns;
      k0 = nk;
      cur = nxt;
    }
  }

  // output row addresses: a row-address table or a plain row stride
  int64_t crow[FM][4];
  if (tk.c_rowtab) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = tl.m0 + wm0 + i * 16 + (lane >> 4) + 4 * r;
        crow[i][r] = slots.crows[tk.c_row + min(row, M - 1)] - tk.c_off;
      }
  } else {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) crow[i][r] = (int64_t)(tl.m0 + wm0 + i * 16 + (lane >> 4) + 4 * r) * tk.ldc;
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
          const int64_t ix = crow[i][r] + col;
          if constexpr (CPLX)
            reinterpret_cast<double2 *>(C)[ix] = double2{acc[0][i][j][r], acc[1][i][j][r]};
          else
            C[ix] = acc[0][i][j][r];
        }
      }
    }
}

template <bool CPLX>
static void launch_shape(int bm, int bn, int bk, const GemmTile *tiles, const GemmTask *tasks, const GemmSeg *segs,
                         const Slots &slots, int ntiles, hipStream_t st) {
  dim3 grid(ntiles), block(256);
#define QLG_SHAPE(M_, N_, K_)                                                                                \
  if (bm == M_ && bn == N_ && bk == K_) {                                                                    \
    hipLaunchKernelGGL((grouped_gemm_f64_kernel<M_, N_, K_, CPLX>), grid, block, 0, st, tiles, tasks, segs,   \
                       slots, ntiles);                                                                       \
    return;                                                                                                  \
  }
  QLG_SHAPE(32, 32, 16)
  QLG_SHAPE(64, 32, 16)
  QLG_SHAPE(32, 64, 16)
  QLG_SHAPE(64, 64, 16)
  QLG_SHAPE(128, 64, 16)
  QLG_SHAPE(64, 128, 16)
  QLG_SHAPE(128, 128, 16)
  if constexpr (!CPLX) {  // two accumulator planes of complex tiles do not fit BK=32 double buffers
    QLG_SHAPE(64, 16, 16)
    QLG_SHAPE(128, 16, 16)
    QLG_SHAPE(32, 32, 32)
    QLG_SHAPE(64, 32, 32)
    QLG_SHAPE(32, 64, 32)
    QLG_SHAPE(64, 64, 32)
    QLG_SHAPE(128, 64, 32)
    QLG_SHAPE(64, 128, 32)
    QLG_SHAPE(64, 16, 32)
    QLG_SHAPE(128, 16, 32)
  }
#undef QLG_SHAPE
  throw Error(QLG_ERR_UNSUPPORTED, "unsupported FP64 tile shape");
}

// Split-K reduction: every thread sums up to four elements of one task's
// partial panels in part order and writes them where the task's epilogue
// would (row table or row stride).
template <bool CPLX>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const RedTask *__restrict__ red,
                                                            const RedChunk *__restrict__ chunks, Slots slots,
                                                            int nchunks) {
  if ((int)blockIdx.x >= nchunks) return;
  if (slots.skip && *slots.skip) return;   // the Krylov solve has stopped
  const RedChunk ch = chunks[blockIdx.x];
  const RedTask t = red[ch.task];
  const int64_t panel = (int64_t)t.M * t.N;
  const double *W = slots.p[SLOT_W];
  double *C = slots.p[t.c_slot];
#pragma unroll
  for (int j = 0; j < kRedChunk / 256; ++j) {
    const int64_t e = (int64_t)ch.e0 + threadIdx.x + j * 256;
    if (e >= panel || e >= (int64_t)ch.e0 + kRedChunk) break;
    const int64_t r = e / t.N, c = e % t.N;
    const int64_t ix = (t.c_rowtab ? slots.crows[t.c_row + r] - t.c_off : r * t.ldc) + c + t.c_off;
    if constexpr (CPLX) {
      double2 acc = reinterpret_cast<const double2 *>(W)[t.wbase + e];
      for (int q = 1; q < t.parts; ++q) {
        const double2 v = reinterpret_cast<const double2 *>(W)[t.wbase + q * panel + e];
        acc.x += v.x;
        acc.y += v.y;
      }
      reinterpret_cast<double2 *>(C)[ix] = acc;
    } else {
      double acc = W[t.wbase + e];
      for (int q = 1; q < t.parts; ++q) acc += W[t.wbase + q * panel + e];
      C[ix] = acc;
    }
  }
}

void launch_splitk_reduce(bool cplx, const RedTask *red, const RedChunk *chunks, int nchunks, const Slots &slots,
                          hipStream_t st) {
  if (nchunks <= 0) return;
  if (cplx)
    hipLaunchKernelGGL(splitk_reduce_kernel<true>, dim3(nchunks), dim3(256), 0, st, red, chunks, slots, nchunks);
  else
    hipLaunchKernelGGL(splitk_reduce_kernel<false>, dim3(nchunks), dim3(256), 0, st, red, chunks, slots, nchunks);
  QLG_HIP(hipGetLastError());
}

int gemm_bk() {
  static int bk = [] {
    const char *v = std::getenv("QLG_BK");
    return (v && std::atoi(v) == 32) ? 32 : 16;
  }();
  return bk;
}

void launch_grouped_gemm_f64(int bm, int bn, bool cplx, const GemmTile *tiles, const GemmTask *tasks,
                             const GemmSeg *segs, const Slots &slots, int ntiles, hipStream_t st) {
  if (ntiles <= 0) return;
  const int bk = (cplx || bm * bn > 64 * 128) ? 16 : gemm_bk();
  if (cplx)
    launch_shape<true>(bm, bn, bk, tiles, tasks, segs, slots, ntiles, st);
  else
    launch_shape<false>(bm, bn, bk, tiles, tasks, segs, slots, ntiles, st);
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
