// CPU baseline: the reference's two-site effective-Hamiltonian matvec on the
// host, as the reference computes it.
//
// TEST AND BASELINE INFRASTRUCTURE ONLY.  The product path never loads this
// library; bench.py times it as the CPU baseline and tests/ check it against
// the numpy oracle.
//
// eff_ham_mul_two_site_state (include/qlmps/algorithm/vmps/lanczos_vmps_solver_impl.h:260-289)
// is four block-sparse contractions:
//   T1 = Contract(L, psi, {{2}, {0}})                        [a', w, s1, s2, b]
//   T2 = Contract<true, true>(T1, W1, 1, 0, 2)               [s2, b, a', s1', y]
//   T3 = Contract<true, true>(T2, W2, 4, 0, 2)               [b, a', s1', s2', z]
//   out = Contract<true, false>(T3, R, 4, 1, 2)              [a', s1', s2', b']
// (rotated contractions: the contracted axes are cyclic runs, the result
// keeps each operand's free axes in cyclic order after its run).  Each
// contraction here works as qlten's does on the host: every operand block is
// brought to matrix form (free x contracted, or contracted x free) by an axis
// permutation into scratch where its layout needs one, and every pair of
// blocks with matching contracted sectors is one dgemm accumulated into its
// output block.  Output blocks are distributed over OpenMP threads (largest
// first), each dgemm single-threaded.  The dgemm comes from a host BLAS
// loaded with dlopen (MKL, a system OpenBLAS, or an OpenBLAS build given by
// path), else from the blocked loop below.  Real FP64 only.
#include <dlfcn.h>
#include <omp.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "../include/qlgpu.h"

namespace {

// ---------------------------------------------------------------------------
// BLAS
// ---------------------------------------------------------------------------
using cblas_dgemm_t = void (*)(int, int, int, int, int, int, double, const double *, int, const double *, int, double,
                               double *, int);
using set_threads_t = void (*)(int);
constexpr int kRowMajor = 101, kNoTrans = 111;

cblas_dgemm_t g_dgemm = nullptr;
std::string g_blas = "none";

// C[M x N] += A[M x K] B[K x N], row-major, leading dimensions K, N, N
void own_dgemm(int M, int N, int K, const double *A, const double *B, double *C) {
  constexpr int BK = 128, BN = 256;
  for (int k0 = 0; k0 < K; k0 += BK)
    for (int n0 = 0; n0 < N; n0 += BN) {
      const int k1 = std::min(K, k0 + BK), n1 = std::min(N, n0 + BN);
      for (int i = 0; i < M; ++i) {
        double *c = C + (int64_t)i * N;
        for (int k = k0; k < k1; ++k) {
          const double a = A[(int64_t)i * K + k];
          const double *b = B + (int64_t)k * N;
          for (int j = n0; j < n1; ++j) c[j] += a * b[j];
        }
      }
    }
}

void gemm_acc(int M, int N, int K, const double *A, const double *B, double *C) {
  if (M == 0 || N == 0 || K == 0) return;
  if (g_dgemm)
    g_dgemm(kRowMajor, kNoTrans, kNoTrans, M, N, K, 1.0, A, K, B, N, 1.0, C, N);
  else
    own_dgemm(M, N, K, A, B, C);
}

// ---------------------------------------------------------------------------
// host block-sparse tensors
// ---------------------------------------------------------------------------
struct Ax {
  std::vector<std::array<int32_t, QLG_MAX_QN_LEN>> qn;
  std::vector<int64_t> deg;
  int dir = 1;
};
using Co = std::array<int32_t, QLG_MAX_RANK>;

struct HT {
  int rank = 0;
  std::vector<Ax> ax;
  std::vector<Co> blk;
  std::vector<int64_t> off;
  int64_t size = 0;
  int64_t blk_elems(size_t b) const {
    int64_t n = 1;
    for (int i = 0; i < rank; ++i) n *= ax[i].deg[blk[b][i]];
    return n;
  }
};

HT parse(const qlg_bsten *d) {
  HT t;
  t.rank = d->rank;
  int64_t qi = 0, di = 0;
  for (int i = 0; i < d->rank; ++i) {
    Ax a;
    a.dir = d->dir[i];
    for (int s = 0; s < d->nsect[i]; ++s) {
      std::array<int32_t, QLG_MAX_QN_LEN> q{};
      for (int c = 0; c < d->qn_len; ++c) q[c] = d->sect_qn[qi++];
      a.qn.push_back(q);
      a.deg.push_back(d->sect_deg[di++]);
    }
    t.ax.push_back(a);
  }
  for (int64_t b = 0; b < d->nblocks; ++b) {
    Co c{};
    for (int i = 0; i < d->rank; ++i) c[i] = d->blk_coords[b * d->rank + i];
    t.blk.push_back(c);
    t.off.push_back(d->blk_offset[b]);
    t.size = std::max(t.size, d->blk_offset[b] + t.blk_elems(t.blk.size() - 1));
  }
  return t;
}

// out[perm-ordered] = in: dst axis i is src axis perm[i] (row-major both)
void permute_block(const double *src, double *dst, int rank, const int64_t *dims, const int *perm) {
  int64_t sstride[QLG_MAX_RANK] = {0}, ddims[QLG_MAX_RANK] = {0}, n = 1, acc = 1;
  for (int i = rank - 1; i >= 0; --i) {
    sstride[i] = acc;
    acc *= dims[i];
  }
  for (int i = 0; i < rank; ++i) {
    ddims[i] = dims[perm[i]];
    n *= ddims[i];
  }
  int64_t idx[QLG_MAX_RANK] = {0};
  int64_t s = 0;
  const int last = rank - 1;
  const int64_t sl = sstride[perm[last]];
  for (int64_t e = 0; e < n;) {
    // innermost dimension in one run
    const int64_t len = ddims[last];
    for (int64_t j = 0; j < len; ++j) dst[e + j] = src[s + j * sl];
    e += len;
    // carry into the outer indexes
    int k = last - 1;
    while (k >= 0) {
      ++idx[k];
      s += sstride[perm[k]];
      if (idx[k] < ddims[k]) break;
      s -= idx[k] * sstride[perm[k]];
      idx[k] = 0;
      --k;
    }
    if (k < 0) break;
  }
}

// one contraction step: C = A x B over A's axes ca and B's axes cb; result
// axes = A's axes af (in that order) then B's axes bf
struct Step {
  int ra = 0, rb = 0;
  std::vector<int> ca, cb, af, bf;
  bool perm_a = false, perm_b = false;   // operand blocks need a permutation to matrix form
  HT out;
  struct Pair {
    int64_t a, b;
  };
  struct Task {
    int64_t out_blk, M, N;
    std::vector<std::pair<Pair, int64_t>> pairs;   // with K
    double cost;
  };
  std::vector<Task> tasks;
  std::vector<int64_t> a_moff, b_moff;   // offsets of the matrix-form copies
  int64_t a_msize = 0, b_msize = 0;
  double flops = 0.0;
};

bool is_identity(const std::vector<int> &p) {
  for (size_t i = 0; i < p.size(); ++i)
    if (p[i] != (int)i) return false;
  return true;
}

Step plan_step(const HT &A, const HT &B, std::vector<int> ca, std::vector<int> cb, std::vector<int> af,
               std::vector<int> bf) {
  Step s;
  s.ra = A.rank;
  s.rb = B.rank;
  s.ca = ca;
  s.cb = cb;
  s.af = af;
  s.bf = bf;
  std::vector<int> pa(af), pb(cb);
  pa.insert(pa.end(), ca.begin(), ca.end());
  pb.insert(pb.end(), bf.begin(), bf.end());
  s.perm_a = !is_identity(pa);
  s.perm_b = !is_identity(pb);
  s.out.rank = (int)(af.size() + bf.size());
  for (int i : af) s.out.ax.push_back(A.ax[i]);
  for (int i : bf) s.out.ax.push_back(B.ax[i]);
  std::map<std::vector<int32_t>, std::vector<int64_t>> bkey;
  for (size_t b = 0; b < B.blk.size(); ++b) {
    std::vector<int32_t> k;
    for (int i : cb) k.push_back(B.blk[b][i]);
    bkey[k].push_back((int64_t)b);
  }
  std::map<Co, std::vector<Step::Pair>> pairs;
  for (size_t a = 0; a < A.blk.size(); ++a) {
    std::vector<int32_t> k;
    for (int i : ca) k.push_back(A.blk[a][i]);
    auto it = bkey.find(k);
    if (it == bkey.end()) continue;
    for (int64_t b : it->second) {
      Co c{};
      int p = 0;
      for (int i : af) c[p++] = A.blk[a][i];
      for (int i : bf) c[p++] = B.blk[b][i];
      pairs[c].push_back({(int64_t)a, b});
    }
  }
  int64_t o = 0;
  for (auto &kv : pairs) {
    s.out.blk.push_back(kv.first);
    s.out.off.push_back(o);
    o += s.out.blk_elems(s.out.blk.size() - 1);
  }
  s.out.size = o;
  size_t ob = 0;
  for (auto &kv : pairs) {
    Step::Task t;
    t.out_blk = (int64_t)ob;
    t.M = 1;
    t.N = 1;
    for (size_t j = 0; j < af.size(); ++j) t.M *= A.ax[af[j]].deg[kv.first[j]];
    for (size_t j = 0; j < bf.size(); ++j) t.N *= B.ax[bf[j]].deg[kv.first[af.size() + j]];
    t.cost = 0.0;
    for (const Step::Pair &pr : kv.second) {
      int64_t K = 1;
      for (int i : ca) K *= A.ax[i].deg[A.blk[pr.a][i]];
      t.pairs.push_back({pr, K});
      t.cost += 2.0 * double(t.M) * double(t.N) * double(K);
    }
    s.flops += t.cost;
    s.tasks.push_back(std::move(t));
    ++ob;
  }
  std::sort(s.tasks.begin(), s.tasks.end(), [](const Step::Task &x, const Step::Task &y) { return x.cost > y.cost; });
  // matrix-form copies
  s.a_moff.assign(A.blk.size(), -1);
  s.b_moff.assign(B.blk.size(), -1);
  if (s.perm_a)
    for (size_t a = 0; a < A.blk.size(); ++a) {
      s.a_moff[a] = s.a_msize;
      s.a_msize += A.blk_elems(a);
    }
  if (s.perm_b)
    for (size_t b = 0; b < B.blk.size(); ++b) {
      s.b_moff[b] = s.b_msize;
      s.b_msize += B.blk_elems(b);
    }
  return s;
}

// bring every block of T to matrix form (axis order `order`) in dst
void to_matrix_form(const HT &T, const double *data, const std::vector<int> &order, const std::vector<int64_t> &moff,
                    double *dst) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t b = 0; b < (int64_t)T.blk.size(); ++b) {
    int64_t dims[QLG_MAX_RANK];
    for (int i = 0; i < T.rank; ++i) dims[i] = T.ax[i].deg[T.blk[b][i]];
    permute_block(data + T.off[b], dst + moff[b], T.rank, dims, order.data());
  }
}

void run_step(const Step &s, const HT &A, const double *a, const HT &B, const double *b, double *out,
              std::vector<double> &abuf, std::vector<double> &bbuf) {
  const double *am = a, *bm = b;
  if (s.perm_a) {
    std::vector<int> order(s.af);
    order.insert(order.end(), s.ca.begin(), s.ca.end());
    abuf.resize(std::max<int64_t>(s.a_msize, 1));
    to_matrix_form(A, a, order, s.a_moff, abuf.data());
    am = abuf.data();
  }
  if (s.perm_b) {
    std::vector<int> order(s.cb);
    order.insert(order.end(), s.bf.begin(), s.bf.end());
    bbuf.resize(std::max<int64_t>(s.b_msize, 1));
    to_matrix_form(B, b, order, s.b_moff, bbuf.data());
    bm = bbuf.data();
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t ti = 0; ti < (int64_t)s.tasks.size(); ++ti) {
    const Step::Task &t = s.tasks[ti];
    double *c = out + s.out.off[t.out_blk];
    std::memset(c, 0, sizeof(double) * t.M * t.N);
    for (const auto &pk : t.pairs) {
      const double *ap = s.perm_a ? am + s.a_moff[pk.first.a] : am + A.off[pk.first.a];
      const double *bp = s.perm_b ? bm + s.b_moff[pk.first.b] : bm + B.off[pk.first.b];
      gemm_acc((int)t.M, (int)t.N, (int)pk.second, ap, bp, c);
    }
  }
}

struct Plan {
  HT L, W1, W2, R, psi;
  HT Rm;   // R's layout in matrix form [(z, b) x b'] (offsets into rmat)
  const double *dL, *dW1, *dW2, *dR;
  Step s1, s2, s3, s4;
  std::vector<double> t1, t2, t3, abuf, bbuf, rmat;
  double flops = 0.0;
};

}  // namespace

extern "C" {

// Load a BLAS: MKL, a system OpenBLAS, else the OpenBLAS at `path` (symbols
// with the given prefix, e.g. "scipy_"); returns 1 with a BLAS, 0 for the
// built-in loop.  name receives what was loaded.
int qlcpu_load_blas(const char *path, const char *prefix, char *name, int n) {
  struct Cand {
    const char *lib, *dgemm, *threads;
  };
  const Cand cands[] = {{"libmkl_rt.so", "cblas_dgemm", "MKL_Set_Num_Threads"},
                        {"libmkl_rt.so.2", "cblas_dgemm", "MKL_Set_Num_Threads"},
                        {"libopenblas.so.0", "cblas_dgemm", "openblas_set_num_threads"},
                        {"libopenblas.so", "cblas_dgemm", "openblas_set_num_threads"}};
  g_dgemm = nullptr;
  g_blas = "built-in blocked loop";
  for (const Cand &c : cands) {
    void *h = dlopen(c.lib, RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    auto f = reinterpret_cast<cblas_dgemm_t>(dlsym(h, c.dgemm));
    if (!f) continue;
    if (auto st = reinterpret_cast<set_threads_t>(dlsym(h, c.threads))) st(1);
    g_dgemm = f;
    g_blas = c.lib;
    break;
  }
  if (!g_dgemm && path && path[0]) {
    if (void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL)) {
      const std::string pre = prefix ? prefix : "";
      auto f = reinterpret_cast<cblas_dgemm_t>(dlsym(h, (pre + "cblas_dgemm").c_str()));
      if (f) {
        if (auto st = reinterpret_cast<set_threads_t>(dlsym(h, (pre + "openblas_set_num_threads").c_str()))) st(1);
        g_dgemm = f;
        g_blas = std::string("OpenBLAS (") + path + ")";
      }
    }
  }
  if (name && n > 0) {
    std::strncpy(name, g_blas.c_str(), n - 1);
    name[n - 1] = 0;
  }
  return g_dgemm ? 1 : 0;
}

int qlcpu_threads() { return omp_get_max_threads(); }

// Plan the four steps for operands L, W1, W2, R (host data, kept by pointer)
// and states with psi's layout; flops = sum 2 M N K of all dgemms.
void *qlcpu_plan_two_site(const qlg_bsten *L, const qlg_bsten *W1, const qlg_bsten *W2, const qlg_bsten *R,
                          const qlg_bsten *psi, double *flops) {
  if (!L || !W1 || !W2 || !R || !psi) return nullptr;
  for (const qlg_bsten *d : {L, W1, W2, R, psi})
    if (d->dtype != QLG_F64) return nullptr;
  auto p = std::make_unique<Plan>();
  p->L = parse(L);
  p->W1 = parse(W1);
  p->W2 = parse(W2);
  p->R = parse(R);
  p->psi = parse(psi);
  p->dL = static_cast<const double *>(L->data);
  p->dW1 = static_cast<const double *>(W1->data);
  p->dW2 = static_cast<const double *>(W2->data);
  p->dR = static_cast<const double *>(R->data);
  p->s1 = plan_step(p->L, p->psi, {2}, {0}, {0, 1}, {1, 2, 3});
  p->s2 = plan_step(p->s1.out, p->W1, {1, 2}, {0, 1}, {3, 4, 0}, {2, 3});
  p->s3 = plan_step(p->s2.out, p->W2, {4, 0}, {0, 1}, {1, 2, 3}, {2, 3});
  p->s4 = plan_step(p->s3.out, p->R, {4, 0}, {1, 2}, {1, 2, 3}, {0});
  p->t1.resize(std::max<int64_t>(p->s1.out.size, 1));
  p->t2.resize(std::max<int64_t>(p->s2.out.size, 1));
  p->t3.resize(std::max<int64_t>(p->s3.out.size, 1));
  // R in matrix form once (it is an operator: the same in every matvec)
  if (p->s4.perm_b) {
    std::vector<int> order(p->s4.cb);
    order.insert(order.end(), p->s4.bf.begin(), p->s4.bf.end());
    p->rmat.resize(std::max<int64_t>(p->s4.b_msize, 1));
    to_matrix_form(p->R, p->dR, order, p->s4.b_moff, p->rmat.data());
    p->Rm = p->R;
    p->Rm.off = p->s4.b_moff;
    p->s4.perm_b = false;
    p->dR = p->rmat.data();
  } else {
    p->Rm = p->R;
  }
  p->flops = p->s1.flops + p->s2.flops + p->s3.flops + p->s4.flops;
  if (flops) *flops = p->flops;
  return p.release();
}

int qlcpu_out_layout(void *h, int64_t *nblocks, int64_t *size, int32_t *coords, int64_t *offsets) {
  if (!h) return -1;
  const HT &o = static_cast<Plan *>(h)->s4.out;
  if (nblocks) *nblocks = (int64_t)o.blk.size();
  if (size) *size = o.size;
  if (coords)
    for (size_t b = 0; b < o.blk.size(); ++b)
      for (int i = 0; i < o.rank; ++i) coords[b * o.rank + i] = o.blk[b][i];
  if (offsets)
    for (size_t b = 0; b < o.blk.size(); ++b) offsets[b] = o.off[b];
  return 0;
}

int qlcpu_matvec(void *h, const double *psi, double *out) {
  if (!h || !psi || !out) return -1;
  Plan &p = *static_cast<Plan *>(h);
  run_step(p.s1, p.L, p.dL, p.psi, psi, p.t1.data(), p.abuf, p.bbuf);
  run_step(p.s2, p.s1.out, p.t1.data(), p.W1, p.dW1, p.t2.data(), p.abuf, p.bbuf);
  run_step(p.s3, p.s2.out, p.t2.data(), p.W2, p.dW2, p.t3.data(), p.abuf, p.bbuf);
  // step 4 with R already in matrix form
  run_step(p.s4, p.s3.out, p.t3.data(), p.Rm, p.dR, out, p.abuf, p.bbuf);
  return 0;
}

void qlcpu_plan_destroy(void *h) { delete static_cast<Plan *>(h); }

}  // extern "C"
