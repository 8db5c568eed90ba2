// Stage-2 kernel of the effective-Hamiltonian matvec: per charge channel
// (a'-sector, b-sector), Y = C . X where C is the small dense combined MPO
// operator of the channel (row-major [K_out x K_in]) and X / Y are sets of
// rows whose elements live in 2-D strided panels of the stage-1 result and
// of the stage-3 operand: position n of a row is (a' = n / Db, b = n % Db)
// at  off + a' * rs + b.  The kernel is the grouped FP64 MFMA GEMM of
// gemm_f64.hip with the B rows gathered and the C rows scattered through
// per-row (off, rs) tables; it reads every X element once and writes every
// Y element once (HBM-bound).
#include <hip/hip_runtime.h>

#include "common.h"
#include "gemm.h"

namespace qlg {

typedef double dbl4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int BK = 16;
constexpr int PAD = 1;
}  // namespace

template <int BM, int BN, bool CPLX>
__global__ __launch_bounds__(256) void mpo_stage_kernel(const GemmTile *__restrict__ tiles,
                                                        const ChanTask *__restrict__ tasks,
                                                        const RowDesc *__restrict__ rows, const double *__restrict__ C,
                                                        const double *__restrict__ X, double *__restrict__ Y,
                                                        int ntiles) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int A_PER = BM * BK / 256;
  constexpr int B_PER = BK * BN / 256;
  constexpr int NP = CPLX ? 2 : 1;
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small");
  static_assert(256 % BN == 0, "a thread keeps one column across its B loads");
  __shared__ double As[2][NP][BK][BM + PAD];
  __shared__ double Bs[2][NP][BK][BN + PAD];
  __shared__ RowDesc rin[kMaxChanIn];
  __shared__ RowDesc rout[BM];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  if ((int)blockIdx.x >= ntiles) return;
  const GemmTile tl = tiles[blockIdx.x];
  if (tl.task < 0) return;
  const ChanTask tk = tasks[tl.task];
  const int M = tk.kout, N = tk.P, K = tk.kin;
  const double *Cm = C + tk.c_off * NP;
  // row descriptors of this channel's inputs and of the tile's outputs
  for (int i = tid; i < K; i += 256) rin[i] = rows[tk.in_rows + i];
  for (int i = tid; i < BM; i += 256) rout[i] = rows[tk.out_rows + min(tl.m0 + i, M - 1)];
  __syncthreads();

  // this thread's fixed column for B loads
  const int gn_b = tl.n0 + tid % BN;
  const bool col_ok = gn_b < N;
  const int64_t ap_b = col_ok ? gn_b / tk.Db : 0, b_b = col_ok ? gn_b % tk.Db : 0;

  dbl4 acc[NP][FM][FN];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[p][i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
  double ra[NP][A_PER], rb[NP][B_PER];

  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int e = tid + i * 256;
      const int m = e / BK, k = e % BK;
      const int gm = tl.m0 + m, gk = k0 + k;
      const bool ok = gm < M && gk < K;
      const int64_t ix = ok ? (int64_t)gm * K + gk : 0;
      if constexpr (CPLX) {
        const double2 v = reinterpret_cast<const double2 *>(Cm)[ix];
        ra[0][i] = ok ? v.x : 0.0;
        ra[1][i] = ok ? v.y : 0.0;
      } else {
        const double v = Cm[ix];
        ra[0][i] = ok ? v : 0.0;
      }
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int e = tid + i * 256;
      const int gk = k0 + e / BN;
      const bool ok = col_ok && gk < K;
      const RowDesc rd = rin[ok ? gk : 0];
      const int64_t ix = ok ? rd.off + ap_b * rd.rs + b_b : rin[0].off;
      if constexpr (CPLX) {
        const double2 v = reinterpret_cast<const double2 *>(X)[ix];
        rb[0][i] = ok ? v.x : 0.0;
        rb[1][i] = ok ? v.y : 0.0;
      } else {
        const double v = X[ix];
        rb[0][i] = ok ? v : 0.0;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int e = tid + i * 256;
#pragma unroll
      for (int p = 0; p < NP; ++p) As[buf][p][e % BK][e / BK] = ra[p][i];
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int e = tid + i * 256;
#pragma unroll
      for (int p = 0; p < NP; ++p) Bs[buf][p][e / BN][e % BN] = rb[p][i];
    }
  };

  int buf = 0;
  load(0);
  store(0);
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) load(k0 + BK);
    const int krem = K - k0;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      if (kk >= krem) break;
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
    }
    if (!more) break;
    store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // scatter: row m -> rout[m - m0], column n -> (n / Db, n % Db)
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int col = tl.n0 + wn0 + j * 16 + (lane & 15);
    if (col >= N) continue;
    const int64_t ap = col / tk.Db, b = col % tk.Db;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lr = wm0 + i * 16 + (lane >> 4) + 4 * r;
        if (tl.m0 + lr < M) {
          const RowDesc rd = rout[lr];
          const int64_t ix = rd.off + ap * rd.rs + b;
          if constexpr (CPLX)
            reinterpret_cast<double2 *>(Y)[ix] = double2{acc[0][i][j][r], acc[1][i][j][r]};
          else
            Y[ix] = acc[0][i][j][r];
        }
      }
  }
}

// 2-D block copies: dst[r * dld + c] = src[r * sld + c] for each descriptor
template <bool CPLX>
__global__ __launch_bounds__(256) void block_copy_kernel(const CopyDesc *__restrict__ d, const double *__restrict__ src,
                                                         double *__restrict__ dst, int n) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const CopyDesc c = d[i];
  const int64_t tot = (int64_t)c.rows * c.cols;
  for (int64_t e = threadIdx.x; e < tot; e += 256) {
    const int64_t r = e / c.cols, col = e % c.cols;
    if constexpr (CPLX)
      reinterpret_cast<double2 *>(dst)[c.dst + r * c.dld + col] =
          reinterpret_cast<const double2 *>(src)[c.src + r * c.sld + col];
    else
      dst[c.dst + r * c.dld + col] = src[c.src + r * c.sld + col];
  }
}

void launch_mpo_stage(int bm, int bn, bool cplx, const GemmTile *tiles, const ChanTask *tasks, const RowDesc *rows,
                      const double *C, const double *X, double *Y, int ntiles, hipStream_t st) {
  if (ntiles <= 0) return;
  dim3 grid(ntiles), block(256);
#define QLG_MPO(M_, N_)                                                                                       \
  if (bm == M_ && bn == N_) {                                                                                 \
    if (cplx)                                                                                                 \
      hipLaunchKernelGGL((mpo_stage_kernel<M_, N_, true>), grid, block, 0, st, tiles, tasks, rows, C, X, Y,   \
                         ntiles);                                                                             \
    else                                                                                                      \
      hipLaunchKernelGGL((mpo_stage_kernel<M_, N_, false>), grid, block, 0, st, tiles, tasks, rows, C, X, Y,  \
                         ntiles);                                                                             \
    QLG_HIP(hipGetLastError());                                                                               \
    return;                                                                                                   \
  }
  QLG_MPO(32, 64)
  QLG_MPO(64, 64)
  QLG_MPO(32, 128)
  QLG_MPO(64, 128)
#undef QLG_MPO
  throw Error(QLG_ERR_UNSUPPORTED, "unsupported MPO-stage tile shape");
}

void launch_block_copy(bool cplx, const CopyDesc *d, int n, const double *src, double *dst, hipStream_t st) {
  if (n <= 0) return;
  if (cplx)
    hipLaunchKernelGGL(block_copy_kernel<true>, dim3(n), dim3(256), 0, st, d, src, dst, n);
  else
    hipLaunchKernelGGL(block_copy_kernel<false>, dim3(n), dim3(256), 0, st, d, src, dst, n);
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
