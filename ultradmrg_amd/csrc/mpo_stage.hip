// Stage-2 kernel of the effective-Hamiltonian matvec: per charge channel
// (a'-sector, b-sector), Y = C . X where C is the small dense combined MPO
// operator of the channel (row-major [K_out x K_in]) and X / Y are sets of
// rows whose elements live in 2-D strided panels of the stage-1 result and
// of the stage-3 operand: position n of a row is (a' = n / Db, b = n % Db)
// at  off + a' * rs + b.
//
// The stage is bound by HBM traffic: every X element is read once per row
// tile (BM >= K_out in the common case, so once) and every Y element is
// written once, and K_in is short (tens).  A workgroup therefore keeps its
// memory traffic in flight in one burst: the four waves split the tile's
// columns, the C chunk (BM x KC) is staged in LDS, and each lane loads the
// X values of its MFMA B fragments for the whole K chunk straight into
// registers before the first MFMA.  Raw buffer loads return zero outside
// the channel (tile_io.h).  Input-row descriptors are staged in LDS,
// kRowWin at a time (channels with many MPO/physical components take several
// windows), and the C and X loads of a chunk are all issued before the first
// wait, so a chunk costs one memory latency.
#include <hip/hip_runtime.h>

#include "common.h"
#include "gemm.h"
#include "tile_io.h"

namespace qlg {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

namespace {
// K chunk held in registers per round.  FP64: 16 rows, which keeps the
// 32x128 kernel at 6 waves per SIMD (32 rows: 4); the channels' K_in is ~16 on
// average, so the extra chunks of the longer channels cost less than the
// lost occupancy did.  FP32 kernels already run at 7-8 waves with 32.
template <typename T>
constexpr int kChunk = sizeof(T) == 8 ? 16 : 32;
constexpr int kRowWin = 256;

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {   // v_mfma_f64_16x16x4: D lane l, reg r -> row l / 16 + 4 r
  typedef dbl4 acc_t;
  __device__ static acc_t run(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
  __device__ static double load(__amdgpu_buffer_rsrc_t rs, uint32_t elem) { return load_f64(rs, elem == kOob ? kOob : elem * 8u); }
};
template <>
struct Mfma<float> {    // v_mfma_f32_16x16x4: D lane l, reg r -> row 4 (l / 16) + r
  typedef flt4 acc_t;
  __device__ static acc_t run(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static float load(__amdgpu_buffer_rsrc_t rs, uint32_t elem) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, elem == kOob ? kOob : elem * 4u, 0, 0));
  }
};
}  // namespace

// T: element type of C, X and Y (double, or float for the reduced-precision
// modes); CPLX: complex data of that element type, stored (re, im).
// OUTP > 0 (real FP32 only): Y is written as OUTP bf16 piece planes of each
// FP32 result x, h = bf16(x), m = bf16(x - h), l = bf16(x - h - m), plane q
// at yps elements after plane q - 1 (the operand of the pre-split stage 3).
template <int BM, int BN, bool CPLX, typename T, int OUTP = 0>
__global__ __launch_bounds__(256) void mpo_stage_kernel(const ChanTile *__restrict__ tiles,
                                                        const RowDesc *__restrict__ rows, const T *__restrict__ C,
                                                        const T *__restrict__ X, T *__restrict__ Y,
                                                        const int *__restrict__ skip, int64_t yps, int ntiles) {
  using MF = Mfma<T>;
  typedef typename MF::acc_t acc_t;
  constexpr int WN = BN / 4;                // columns per wave
  constexpr int FM = BM / 16, FN = WN / 16;
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int KC = kChunk<T>;
  constexpr int KS = KC / 4;                // MFMA k-steps per chunk
  static_assert(FM >= 1 && FN >= 1, "tile too small");
  __shared__ T As[NP][KC][BM + 1];
  __shared__ RowDesc rin[kRowWin];
  __shared__ RowDesc rout[256];   // BM used; one slot per thread, written unconditionally

  static_assert(kRowWin == 256 && BM <= 256, "one descriptor per thread");
  static_assert(kRowWin % KC == 0, "a descriptor window holds whole K chunks");
  constexpr int CE = BM * KC / 256;         // C-chunk elements per thread
  static_assert(CE * 256 == BM * KC, "C chunk splits evenly over the threads");

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  if ((int)blockIdx.x >= ntiles) return;
  const ChanTile tl = tiles[blockIdx.x];   // issued before the flag load, waited for together
  if (skip && *skip) return;   // the Krylov solve has stopped
  if (tl.m0 < 0) return;
  const ChanTask &tk = tl.c;
  const int M = tk.kout, N = tk.P, K = tk.kin;
  const T *Cm = C + tk.c_off * NP;
  const __amdgpu_buffer_rsrc_t xr = make_rsrc(X);
  const RowDesc *gin = rows + tk.in_rows;
  // Input-row descriptors are only ever read from LDS: rows [w, w + kRowWin)
  // of the channel sit in rin while chunk kc (w = kc rounded down to
  // kRowWin) is loaded.  (A descriptor read that could come from either LDS
  // or memory compiles to a generic load, which waits for every X load
  // issued before it: one memory latency per k-step.)  Both staging loads
  // are issued before either LDS write.
  {
    // clamped indexes: every thread loads an existing descriptor (K, M >= 1)
    const RowDesc din = gin[min(tid, K - 1)];
    const RowDesc dout = rows[tk.out_rows + min(tl.m0 + min(tid, BM - 1), M - 1)];
    // unconditional writes: a conditional one lets the compiler sink its
    // load into the branch and wait for each load in turn
    rin[tid] = din;
    rout[tid] = dout;
  }

  // this lane's columns: (a', b) positions of its B fragments
  int apj[FN], bj[FN];
  bool okj[FN];
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int col = tl.n0 + wave * WN + j * 16 + (lane & 15);
    okj[j] = col < N;
    apj[j] = okj[j] ? col / tk.Db : 0;
    bj[j] = okj[j] ? col % tk.Db : 0;
  }

  acc_t acc[NP][FM][FN];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[p][i][j] = acc_t{0, 0, 0, 0};

  for (int kc = 0; kc < K; kc += KC) {
    __syncthreads();  // descriptors staged / previous chunk's As and rin reads done
    const int win = kc & ~(kRowWin - 1);
    if (kc > 0 && kc == win) {  // channels with more than kRowWin input rows: the next window
      if (tid < min(K - win, kRowWin)) rin[tid] = gin[win + tid];
      __syncthreads();
    }
    // the C chunk (rows m0.., columns kc..), loaded into registers first and
    // written to LDS after the X loads below are issued, so that the chunk
    // waits on memory once
    T cv[NP][CE];
#pragma unroll
    for (int q = 0; q < CE; ++q) {
      const int e = tid + 256 * q;
      const int m = e / KC, k = e % KC;
      const int gm = tl.m0 + m, gk = kc + k;
      const bool ok = gm < M && gk < K;
      const int64_t ix = ok ? (int64_t)gm * K + gk : 0;
      if constexpr (CPLX) {
        if constexpr (sizeof(T) == 8) {
          const double2 v = reinterpret_cast<const double2 *>(Cm)[ix];
          cv[0][q] = ok ? v.x : T(0);
          cv[1][q] = ok ? v.y : T(0);
        } else {
          const float2 v = reinterpret_cast<const float2 *>(Cm)[ix];
          cv[0][q] = ok ? v.x : T(0);
          cv[1][q] = ok ? v.y : T(0);
        }
      } else {
        cv[0][q] = ok ? Cm[ix] : T(0);
      }
    }
    // B fragments of the whole chunk: lane holds X[row kc + 4 s + lane / 16][its column j].
    // k-steps that start at or past K are skipped (wave-uniform), which saves
    // their load instructions and MFMAs: K_in is short, so most of a chunk's
    // steps are empty
    T rb[NP][KS][FN];
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      if (kc + 4 * st >= K) break;
      const int gk = kc + 4 * st + (lane >> 4);
      const bool kok = gk < K;
      const RowDesc rd = rin[kok ? gk - win : 0];
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const uint32_t el = (kok && okj[j]) ? (uint32_t)((int)rd.off + apj[j] * (int)rd.rs + bj[j]) : kOob;
        if constexpr (CPLX && sizeof(T) == 8) {
          const double2 v = load_c128(xr, el == kOob ? kOob : el * 16u);
          rb[0][st][j] = v.x;
          rb[1][st][j] = v.y;
        } else if constexpr (CPLX) {
          const float2 v = load_c64(xr, el == kOob ? kOob : el * 8u);
          rb[0][st][j] = v.x;
          rb[1][st][j] = v.y;
        } else {
          rb[0][st][j] = MF::load(xr, el);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < CE; ++q) {
      const int e = tid + 256 * q;
#pragma unroll
      for (int p = 0; p < NP; ++p) As[p][e % KC][e / KC] = cv[p][q];
    }
    __syncthreads();  // C chunk staged
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      if (kc + 4 * st >= K) break;
      T af[NP][FM];
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int i = 0; i < FM; ++i) af[p][i] = As[p][4 * st + (lane >> 4)][i * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          acc[0][i][j] = MF::run(af[0][i], rb[0][st][j], acc[0][i][j]);
          if constexpr (CPLX) {
            acc[0][i][j] = MF::run(-af[1][i], rb[1][st][j], acc[0][i][j]);
            acc[1][i][j] = MF::run(af[0][i], rb[1][st][j], acc[1][i][j]);
            acc[1][i][j] = MF::run(af[1][i], rb[0][st][j], acc[1][i][j]);
          }
        }
    }
  }
  // scatter: row m -> rout[m - m0], column n -> (n / Db, n % Db).  The
  // lane's row descriptors are all read from LDS before the first store
  // (every rout slot holds a valid descriptor): read inside the guarded
  // store, each would wait for its own LDS round trip.
  // (row strides are leading dimensions of the Y panels, below 2^31)
  int64_t ro[FM][4];
  int32_t rr[FM][4];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const RowDesc d = rout[i * 16 + MF::row(lane, r)];
      ro[i][r] = d.off;
      rr[i][r] = (int32_t)d.rs;
    }
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    if (!okj[j]) continue;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lr = i * 16 + MF::row(lane, r);
        if (tl.m0 + lr < M) {
          const int64_t ix = ro[i][r] + (int64_t)apj[j] * rr[i][r] + bj[j];
          if constexpr (OUTP > 0) {
            float x = acc[0][i][j][r];
#pragma unroll
            for (int q = 0; q < OUTP; ++q) {
              const __bf16 h = (__bf16)x;
              reinterpret_cast<__bf16 *>(Y)[q * yps + ix] = h;
              x -= (float)h;
            }
          } else if constexpr (CPLX && sizeof(T) == 8)
            reinterpret_cast<double2 *>(Y)[ix] = double2{acc[0][i][j][r], acc[1][i][j][r]};
          else if constexpr (CPLX)
            reinterpret_cast<float2 *>(Y)[ix] = float2{acc[0][i][j][r], acc[1][i][j][r]};
          else
            Y[ix] = acc[0][i][j][r];
        }
      }
  }
}

// 2-D block copies: dst[r * dld + c] = src[r * sld + c] for each descriptor.
// Workgroup (x, y) copies rows y, y + kCopySplit, ... of descriptor x; each
// wave takes one row at a time and its lanes walk the columns, so loads and
// stores are contiguous runs and no per-element division is needed.
constexpr int kCopySplit = 8;

template <bool CPLX>
__global__ __launch_bounds__(256) void block_copy_kernel(const CopyDesc *__restrict__ d, const double *__restrict__ src,
                                                         double *__restrict__ dst, int n) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const CopyDesc c = d[i];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = blockIdx.y + kCopySplit * wave; r < c.rows; r += kCopySplit * 4) {
    const int64_t so = c.src + (int64_t)r * c.sld, dof = c.dst + (int64_t)r * c.dld;
    for (int col = lane; col < c.cols; col += 64) {
      if constexpr (CPLX)
        reinterpret_cast<double2 *>(dst)[dof + col] = reinterpret_cast<const double2 *>(src)[so + col];
      else
        dst[dof + col] = src[so + col];
    }
  }
}

// transposed copies (stacking block operators): one workgroup per descriptor
template <bool CPLX>
__global__ __launch_bounds__(256) void trans_copy_kernel(const TransDesc *__restrict__ d, const double *__restrict__ src,
                                                         double *__restrict__ dst, int n) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const TransDesc c = d[i];
  const int64_t tot = (int64_t)c.nz * c.nx;
  for (int64_t e = threadIdx.x; e < tot; e += 256) {
    const int64_t x = e / c.nz, z = e % c.nz;   // consecutive threads write consecutive z
    if constexpr (CPLX)
      reinterpret_cast<double2 *>(dst)[c.dst + x * c.dsx + z] =
          reinterpret_cast<const double2 *>(src)[c.src + z * c.sld + x];
    else
      dst[c.dst + x * c.dsx + z] = src[c.src + z * c.sld + x];
  }
}

// the same copies with a source buffer per descriptor (many operators, one launch)
template <bool CPLX>
__global__ __launch_bounds__(256) void trans_copy_multi_kernel(const TransDesc *__restrict__ d,
                                                               const double *const *__restrict__ srcs,
                                                               double *__restrict__ dst, int n) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const TransDesc c = d[i];
  const double *src = srcs[i];
  const int64_t tot = (int64_t)c.nz * c.nx;
  for (int64_t e = threadIdx.x; e < tot; e += 256) {
    const int64_t x = e / c.nz, z = e % c.nz;
    if constexpr (CPLX)
      reinterpret_cast<double2 *>(dst)[c.dst + x * c.dsx + z] =
          reinterpret_cast<const double2 *>(src)[c.src + z * c.sld + x];
    else
      dst[c.dst + x * c.dsx + z] = src[c.src + z * c.sld + x];
  }
}

void launch_trans_copy_multi(bool cplx, const TransDesc *d, const double *const *srcs, int n, double *dst,
                             hipStream_t st) {
  if (n <= 0) return;
  if (cplx)
    hipLaunchKernelGGL(trans_copy_multi_kernel<true>, dim3(n), dim3(256), 0, st, d, srcs, dst, n);
  else
    hipLaunchKernelGGL(trans_copy_multi_kernel<false>, dim3(n), dim3(256), 0, st, d, srcs, dst, n);
  QLG_HIP(hipGetLastError());
}

void launch_trans_copy(bool cplx, const TransDesc *d, int n, const double *src, double *dst, hipStream_t st) {
  if (n <= 0) return;
  if (cplx)
    hipLaunchKernelGGL(trans_copy_kernel<true>, dim3(n), dim3(256), 0, st, d, src, dst, n);
  else
    hipLaunchKernelGGL(trans_copy_kernel<false>, dim3(n), dim3(256), 0, st, d, src, dst, n);
  QLG_HIP(hipGetLastError());
}

void launch_mpo_stage(int bm, int bn, bool cplx, bool f32, int out_planes, int64_t yps, const ChanTile *tiles,
                      const RowDesc *rows, const void *C, const void *X, void *Y, const int *skip, int ntiles,
                      hipStream_t st) {
  if (ntiles <= 0) return;
  dim3 grid(ntiles), block(256);
  const double *Cd = static_cast<const double *>(C), *Xd = static_cast<const double *>(X);
  const float *Cf = static_cast<const float *>(C), *Xf = static_cast<const float *>(X);
  double *Yd = static_cast<double *>(Y);
  float *Yf = static_cast<float *>(Y);
#define QLG_MPO(M_, N_)                                                                                       \
  if (bm == M_ && bn == N_) {                                                                                 \
    if (out_planes == 3 && f32 && !cplx)                                                                      \
      hipLaunchKernelGGL((mpo_stage_kernel<M_, N_, false, float, 3>), grid, block, 0, st, tiles, rows,  \
                         Cf, Xf, Yf, skip, yps, ntiles);                                                      \
    else if (out_planes == 2 && f32 && !cplx)                                                                 \
      hipLaunchKernelGGL((mpo_stage_kernel<M_, N_, false, float, 2>), grid, block, 0, st, tiles, rows,  \
                         Cf, Xf, Yf, skip, yps, ntiles);                                                      \
    else if (out_planes != 0)                                                                                 \
      throw Error(QLG_ERR_UNSUPPORTED, "MPO stage: piece-plane output is for real FP32 data");               \
    else if (f32 && cplx)                                                                                     \
      hipLaunchKernelGGL((mpo_stage_kernel<M_, N_, true, float>), grid, block, 0, st, tiles, rows, Cf,  \
                         Xf, Yf, skip, 0, ntiles);                                                                     \
    else if (f32)                                                                                             \
      hipLaunchKernelGGL((mpo_stage_kernel<M_, N_, false, float>), grid, block, 0, st, tiles, rows, Cf, \
                         Xf, Yf, skip, 0, ntiles);                                                                     \
    else if (cplx)                                                                                            \
      hipLaunchKernelGGL((mpo_stage_kernel<M_, N_, true, double>), grid, block, 0, st, tiles, rows, Cd, \
                         Xd, Yd, skip, 0, ntiles);                                                                     \
    else                                                                                                      \
      hipLaunchKernelGGL((mpo_stage_kernel<M_, N_, false, double>), grid, block, 0, st, tiles, rows,   \
                         Cd, Xd, Yd, skip, 0, ntiles);                                                                 \
    QLG_HIP(hipGetLastError());                                                                               \
    return;                                                                                                   \
  }
  QLG_MPO(16, 128)
  QLG_MPO(32, 128)
  QLG_MPO(64, 128)
  QLG_MPO(16, 256)
  QLG_MPO(32, 256)
  QLG_MPO(64, 64)
#undef QLG_MPO
  throw Error(QLG_ERR_UNSUPPORTED, "unsupported MPO-stage tile shape");
}

void launch_block_copy(bool cplx, const CopyDesc *d, int n, const double *src, double *dst, hipStream_t st) {
  if (n <= 0) return;
  if (cplx)
    hipLaunchKernelGGL(block_copy_kernel<true>, dim3(n, kCopySplit), dim3(256), 0, st, d, src, dst, n);
  else
    hipLaunchKernelGGL(block_copy_kernel<false>, dim3(n, kCopySplit), dim3(256), 0, st, d, src, dst, n);
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
