// Grouped, segmented GEMM task lists shared by the planner and the kernels.
//
// A task computes one output panel C[M x N] (row-major, ldc) as the sum over
// its K-segments of A_s[M x K_s] * op(B_s), where A_s is row-major (lda) and
// op(B_s) is B_s[K_s x N] (row-major, ldb) or B_s^T with B_s[N x K_s]
// (row-major, ldb).  Segments are summed in list order inside one workgroup,
// so results are deterministic and no atomics are used.  Output panels are
// overwritten.
//
// Operand addresses are (base slot, element offset) pairs; the launcher
// passes the slot pointers by value, so a task list can be reused for every
// matvec whatever buffers the Krylov vectors live in.  Slot 0 is the null
// pointer: an offset in slot 0 is an absolute element address.
#pragma once
#include <cstdint>

namespace qlg {

enum BaseSlot : uint8_t {
  SLOT_ABS = 0,
  SLOT_IN = 1,    // Krylov input vector
  SLOT_OUT = 2,   // Krylov output vector
  SLOT_L = 3,     // left environment
  SLOT_R = 4,     // right environment
  SLOT_X = 5,     // first intermediate
  SLOT_Y = 6,     // second intermediate
  SLOT_C = 7,     // combined MPO coefficient matrices
  SLOT_W = 8,     // split-K partial panels
  NSLOTS = 9
};

// Segment flags: op(A) = A^T when A is stored [K x M] (row-major, lda);
// op(B) = B^T when B is stored [N x K]; conjugation applies to complex data.
enum SegFlags : uint8_t { SEG_TRANS_B = 1, SEG_TRANS_A = 2, SEG_CONJ_A = 4, SEG_CONJ_B = 8 };

struct GemmSeg {
  int64_t a_off, b_off;
  int64_t lda, ldb;
  int32_t K;
  uint8_t a_slot, b_slot, flags, pad;
};

// c_rowtab != 0: output row r of the task is written at element
// crow_table[c_row + r] of the C slot instead of c_off + r * ldc.
struct GemmTask {
  int64_t c_off;
  int64_t ldc;
  int32_t M, N;
  int32_t seg_begin, seg_count;
  int32_t c_row;
  uint8_t c_slot, c_rowtab, pad[2];
};

// Stage-2 (MPO) channel: K_out x K_in dense operator at c_off in the
// coefficient buffer; its input / output rows are RowDesc entries starting
// at in_rows / out_rows; positions n = a' * Db + b for n < P.
struct ChanTask {
  int64_t c_off;
  int32_t kin, kout;
  int32_t P, Db;
  int32_t in_rows, out_rows;
};


// An MPO-stage tile with its channel inlined, so that a workgroup's first
// memory access is one scalar load instead of tile -> channel; m0 < 0 marks
// an empty launch slot.
struct ChanTile {
  ChanTask c;
  int32_t m0, n0;
  int32_t pad[2];
};

struct RowDesc {
  int64_t off;   // element offset of position (0, 0)
  int64_t rs;    // element stride between consecutive a'
};

// 2-D block copy (element offsets and leading dimensions)
struct CopyDesc {
  int64_t src, dst;
  int64_t sld, dld;
  int32_t rows, cols;
};

// transposed 2-D copy: dst[x * dsx + z] = src[z * sld + x] for z < nz, x < nx
struct TransDesc {
  int64_t src, dst;
  int64_t sld, dsx;
  int32_t nz, nx;
};

// Split-K: a task whose K is long runs as `parts` tasks over consecutive K
// ranges, each writing a dense partial panel [M x N] (ldc = N) at
// wbase + part * M * N in the SLOT_W workspace; the reduce kernel sums the
// panels in part order (deterministic) and writes the task's output exactly
// as the task's own epilogue would (row table or row stride).
struct RedTask {
  int64_t c_off, ldc, wbase;
  int32_t M, N, parts, c_row;
  uint8_t c_slot, c_rowtab, pad[6];
};

// reduce launch unit: elements [e0, e0 + kRedChunk) of task `task`'s M x N panel
struct RedChunk {
  int32_t task, e0;
};
constexpr int kRedChunk = 1024;

struct GemmTile {
  int32_t task;
  int32_t m0, n0;
  int32_t pad;
};

struct Slots {
  double *p[NSLOTS];
  // Operands held as bf16 piece planes (the split-BF16 modes): plane q of the
  // slot's operand starts ps[slot] elements after plane q - 1.
  int64_t ps[NSLOTS];
  const int64_t *crows;   // row-address table of tasks with c_rowtab set
  // Device flag of a Krylov solve that runs ahead of the host: when it is set
  // (the solve has stopped), every launch of the matvec returns at once.
  // Null: always run.
  const int *skip;
};

}  // namespace qlg
