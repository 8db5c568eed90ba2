// The reference's Lanczos known-answer test (tests/test_algorithm/test_lanczos_solver.cc:114-167)
// written against the header-only layer: same operators (glibc rand after srand(0),
// tests/testing_utils.h:26-41), same call LanczosSolver(eff_ham, init, &eff_ham_mul_two_site_state, params).
// Prints "E <energy> ITERS <n>" per case; the pytest driver compares with dense eigenvalues.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "qlmps_amd.h"

using namespace qlmps_amd;
using DTen = BSTensor<QLTEN_Double>;

static double Rand() { return double(rand()) / RAND_MAX; }

static void RandRealSymMat(double *mat, long dim) {
  srand(0);
  for (long i = 0; i < dim; ++i)
    for (long j = 0; j < dim; ++j) mat[i * dim + j] = Rand();
  srand(0);
  for (long i = 0; i < dim; ++i)
    for (long j = 0; j < dim; ++j) mat[j * dim + i] += Rand();
}

int main(int argc, char **argv) {
  const size_t d = 2, D = 20, dh = 2;
  const std::vector<int32_t> q0 = {0};
  Index din({{q0, d}}, IN), dout = InverseIndex(din);
  Index Din({{q0, D}}, IN), Dout = InverseIndex(Din);
  Index vin({{q0, dh}}, IN), vout = InverseIndex(vin);

  std::vector<double> M(D * D), m(d * d);
  RandRealSymMat(M.data(), D);
  RandRealSymMat(m.data(), d);
  std::vector<double> lb(D * dh * D), rb(D * dh * D), sb(dh * d * d * dh, 0.0);
  for (size_t i = 0; i < D; ++i)
    for (size_t j = 0; j < D; ++j)
      for (size_t k = 0; k < dh; ++k) {
        lb[(i * dh + k) * D + j] = M[i * D + j];   // lblock({i, k, j})
        rb[(j * dh + k) * D + i] = M[i * D + j];   // rblock({j, k, i})
      }
  for (size_t i = 0; i < d; ++i)
    for (size_t j = 0; j < d; ++j)
      for (size_t k = 0; k < dh; ++k) sb[((k * d + i) * d + j) * dh + k] = m[i * d + j];  // lsite({k, i, j, k})

  DTen lblock({Din, vout, Dout}), lsite({vin, din, dout, vout}), rblock({Dout, vin, Din});
  lblock.SetBlock({0, 0, 0}, lb);
  lblock.Upload();
  rblock.SetBlock({0, 0, 0}, rb);
  rblock.Upload();
  lsite.SetBlock({0, 0, 0, 0}, sb);
  lsite.Upload();
  DTen rsite(lsite);

  // the two stop rules of the reference test
  const LanczosParams cases[2] = {LanczosParams(1.0E-9), LanczosParams(1.0E-16, 20)};
  for (const auto &params : cases) {
    auto *init = new DTen({Din, dout, dout, Dout});
    std::vector<double> v(D * d * d * D);
    srand(12345);
    for (auto &x : v) x = Rand() - 0.5;
    init->SetBlock({0, 0, 0, 0}, v);
    init->Upload();
    auto res = LanczosSolver<DTen>({&lblock, &lsite, &rsite, &rblock}, init, &eff_ham_mul_two_site_state<double>,
                                   params);
    std::printf("E %.15f ITERS %zu NORM %.15f\n", res.gs_eng, res.iters, res.gs_vec->Norm());
    delete res.gs_vec;
  }
  // a matvec of the caller's own (here: a trivial one) cannot run in the library: refused,
  // and the initial state stays with the caller
  {
    auto *init = new DTen({Din, dout, dout, Dout});
    DTen *(*foreign)(const std::vector<DTen *> &, DTen *) = [](const std::vector<DTen *> &, DTen *s) {
      return new DTen(*s);
    };
    bool refused = false;
    try {
      (void)LanczosSolver<DTen>({&lblock, &lsite, &rsite, &rblock}, init, foreign, LanczosParams(1.0E-9));
    } catch (const std::invalid_argument &) {
      refused = true;
    }
    bool mismatched = false;   // the one-site matvec with four operators
    try {
      (void)LanczosSolver<DTen>({&lblock, &lsite, &rsite, &rblock}, init, &eff_ham_mul_single_site_state<double>,
                                LanczosParams(1.0E-9));
    } catch (const std::invalid_argument &) {
      mismatched = true;
    }
    delete init;
    std::printf("FOREIGN %d %d\n", (int)refused, (int)mismatched);
  }
  // the same operator in the DMRG (super-block) form of lanczos_dmrg_solver_impl.h:61-66:
  // one term group per MPO bond value k, {(L_k, m)} with {(m, R_k)}, L_k[a, a'] = lblock(a', k, a)
  // = M[a'][a] and R_k[b, b'] = rblock(b', k, b) = M[b][b']
  {
    std::vector<DTen> Lk, Rk;
    for (size_t k = 0; k < dh; ++k) {
      Lk.emplace_back(std::vector<Index>{Dout, Din});
      Rk.emplace_back(std::vector<Index>{Din, Dout});
    }
    std::vector<double> Mt(D * D);
    for (size_t a = 0; a < D; ++a)
      for (size_t b = 0; b < D; ++b) Mt[a * D + b] = M[b * D + a];
    for (size_t k = 0; k < dh; ++k) {
      Lk[k].SetBlock({0, 0}, Mt);
      Lk[k].Upload();
      Rk[k].SetBlock({0, 0}, M);
      Rk[k].Upload();
    }
    DTen msite({din, dout});
    msite.SetBlock({0, 0}, m);
    msite.Upload();
    SuperBlockHamiltonianTerms<DTen> terms;
    for (size_t k = 0; k < dh; ++k)
      terms.emplace_back(BlockSiteHamiltonianTermGroup<DTen>{{&Lk[k], &msite}},
                         SiteBlockHamiltonianTermGroup<DTen>{{&msite, &Rk[k]}});
    auto *init = new DTen({Din, dout, dout, Dout});
    std::vector<double> v(D * d * d * D);
    srand(12345);
    for (auto &x : v) x = Rand() - 0.5;
    init->SetBlock({0, 0, 0, 0}, v);
    init->Upload();
    auto res = LanczosSolver<DTen>(terms, init, LanczosParams(1.0E-9));
    std::printf("SB %.15f ITERS %zu NORM %.15f\n", res.gs_eng, res.iters, res.gs_vec->Norm());
    delete res.gs_vec;
  }
  // tridiagonal helpers with the reference signatures
  double e;
  double *vec = nullptr;
  TridiagGsSolver({2.0, 1.0, 3.0}, {0.5, 0.25}, 3, e, vec, 'V');
  std::printf("TRIDIAG %.15f %.15f\n", e, vec[0]);
  delete[] vec;
  // the plans above took their workspaces from the context's block cache
  auto &ctx = qlmps_amd::Context::Default();
  const auto st = ctx.MemoryStats();
  ctx.ReleaseCached();
  std::printf("CACHE %lld %lld %lld\n", (long long)st[1], (long long)st[2], (long long)ctx.MemoryStats()[0]);
  return 0;
}
