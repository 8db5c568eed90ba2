/*
 * qlmps_amd.h -- header-only C++ host layer over the qlgpu C ABI.
 *
 * It restores the call surface of the reference's Krylov solvers so that a
 * caller written against them keeps its code:
 *   LanczosParams            include/qlmps/algorithm/lanczos_params.h:27-46
 *   LanczosRes, LanczosSolver, eff_ham_mul_two_site_state,
 *   eff_ham_mul_single_site_state, TridiagGsSolver
 *                            include/qlmps/algorithm/vmps/lanczos_vmps_solver_impl.h:98-355
 *   ExpmvRes, LanczosExpmvSolver, TridiagExpme1Solver, Distance
 *                            include/qlmps/algorithm/tdvp/lanczos_expmv_solver_impl.h:41-269
 * with the same ownership rules: the initial state passed to a solver is
 * consumed (the solver deletes it), and the returned gs_vec / expmv tensor is
 * newly allocated and owned by the caller.  Failures throw std::runtime_error
 * with the library's message (the reference exit()s; nothing here does).
 *
 * The tensor type is qlmps_amd::BSTensor<Elem>: quantum-number bookkeeping on
 * the host, one flat device buffer in the canonical block layout (blocks in
 * lexicographic order of sector coordinates, dense row-major, offsets = prefix
 * sums).  Only the C ABI is used; link with -lqlgpu.
 */
#ifndef QLMPS_AMD_H
#define QLMPS_AMD_H

#include <algorithm>
#include <array>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <utility>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "qlgpu.h"

namespace qlmps_amd {

using QLTEN_Double = double;
using QLTEN_Complex = std::complex<double>;

enum TenIndexDirType { IN = -1, OUT = 1 };

inline void Check(int rc, qlg_ctx ctx = nullptr) {
  if (rc != QLG_OK) throw std::runtime_error(std::string("qlgpu error ") + std::to_string(rc) + ": " +
                                             qlg_last_error(ctx));
}

// ---------------------------------------------------------------------------
// context (one per device, created on first use)
// ---------------------------------------------------------------------------
class Context {
 public:
  explicit Context(int device = 0) { Check(qlg_ctx_create(device, nullptr, &h_)); }
  ~Context() { qlg_ctx_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  qlg_ctx get() const { return h_; }
  // the device block cache of plan workspaces: return it to the driver, or
  // read {cached bytes, hipMalloc calls, reuses, hipFree calls}
  void ReleaseCached() { Check(qlg_ctx_release_cached(h_), h_); }
  std::array<int64_t, 4> MemoryStats() const {
    std::array<int64_t, 4> st{};
    Check(qlg_ctx_memory_stats(h_, st.data()), h_);
    return st;
  }
  static Context &Default() {
    static Context c(0);
    return c;
  }

 private:
  qlg_ctx h_ = nullptr;
};

// ---------------------------------------------------------------------------
// quantum numbers and indexes
// ---------------------------------------------------------------------------
struct QNSector {
  std::vector<int32_t> qn;  // 1 (U(1)) or 2 (U(1)xU(1)) charges
  size_t deg;
};

class Index {
 public:
  Index() = default;
  Index(std::vector<QNSector> sectors, TenIndexDirType dir) : sectors_(std::move(sectors)), dir_(dir) {}
  const std::vector<QNSector> &sectors() const { return sectors_; }
  TenIndexDirType GetDir() const { return dir_; }
  size_t dim() const {
    size_t d = 0;
    for (auto &s : sectors_) d += s.deg;
    return d;
  }
  bool operator==(const Index &o) const {
    if (dir_ != o.dir_ || sectors_.size() != o.sectors_.size()) return false;
    for (size_t i = 0; i < sectors_.size(); ++i)
      if (sectors_[i].qn != o.sectors_[i].qn || sectors_[i].deg != o.sectors_[i].deg) return false;
    return true;
  }

 private:
  std::vector<QNSector> sectors_;
  TenIndexDirType dir_ = OUT;
};

inline Index InverseIndex(const Index &i) {
  return Index(i.sectors(), i.GetDir() == IN ? OUT : IN);
}

template <typename Elem>
constexpr int DtypeOf() {
  return std::is_same<Elem, QLTEN_Complex>::value ? QLG_C128 : QLG_F64;
}

// ---------------------------------------------------------------------------
// block-sparse tensor with a device buffer
// ---------------------------------------------------------------------------
template <typename Elem>
class BSTensor {
 public:
  using Coords = std::vector<int32_t>;

  BSTensor() = default;
  explicit BSTensor(std::vector<Index> idx, Context &ctx = Context::Default()) : idx_(std::move(idx)), ctx_(&ctx) {}
  ~BSTensor() { Release(); }
  BSTensor(const BSTensor &o) : idx_(o.idx_), blk_(o.blk_), off_(o.off_), host_(o.host_), ctx_(o.ctx_) {
    if (o.dev_) {
      Alloc(o.size_);
      Check(qlg_memcpy_d2d(ctx_->get(), dev_, o.dev_, o.size_ * sizeof(Elem)), ctx_->get());
    }
  }
  BSTensor(BSTensor &&o) noexcept { Swap(o); }
  BSTensor &operator=(BSTensor o) {
    Swap(o);
    return *this;
  }

  // ----- building (host side, then Upload) ---------------------------------
  // Set a whole block (dense row-major over the block's degeneracies).
  void SetBlock(const Coords &c, std::vector<Elem> data) {
    if (data.size() != BlockSize(c)) throw std::runtime_error("SetBlock: wrong block size");
    host_[c] = std::move(data);
  }
  // Move the host blocks into one device buffer in canonical order.
  void Upload() {
    blk_.clear();
    off_.clear();
    std::vector<Elem> flat;
    for (auto &kv : host_) {
      blk_.push_back(kv.first);
      off_.push_back((int64_t)flat.size());
      flat.insert(flat.end(), kv.second.begin(), kv.second.end());
    }
    Alloc(flat.size());
    Check(qlg_memcpy_h2d(ctx_->get(), dev_, flat.data(), flat.size() * sizeof(Elem)), ctx_->get());
    host_.clear();
  }
  // Adopt an existing layout with a fresh (uninitialised) device buffer.
  void SetLayout(std::vector<Coords> blk, std::vector<int64_t> off, size_t total) {
    blk_ = std::move(blk);
    off_ = std::move(off);
    Alloc(total);
  }
  std::map<Coords, std::vector<Elem>> Download() const {
    std::vector<Elem> flat(size_);
    if (size_) Check(qlg_memcpy_d2h(ctx_->get(), flat.data(), dev_, size_ * sizeof(Elem)), ctx_->get());
    std::map<Coords, std::vector<Elem>> out;
    for (size_t b = 0; b < blk_.size(); ++b) {
      const size_t n = BlockSize(blk_[b]);
      out[blk_[b]] = std::vector<Elem>(flat.begin() + off_[b], flat.begin() + off_[b] + n);
    }
    return out;
  }

  // ----- reference-named queries ---------------------------------------------
  size_t Rank() const { return idx_.size(); }
  size_t size() const { return size_; }     // number of stored elements
  const std::vector<Index> &GetIndexes() const { return idx_; }
  const std::vector<Coords> &Blocks() const { return blk_; }
  const std::vector<int64_t> &Offsets() const { return off_; }
  Elem *DeviceData() { return dev_; }
  const Elem *DeviceData() const { return dev_; }
  Context &ctx() const { return *ctx_; }
  size_t BlockSize(const Coords &c) const {
    size_t n = 1;
    for (size_t i = 0; i < idx_.size(); ++i) n *= idx_[i].sectors()[c[i]].deg;
    return n;
  }

  // 2-norm, and in-place normalisation returning the old norm
  // (qlten Normalize / QuasiNormalize for bosonic tensors)
  double Norm() const {
    double n = 0;
    Check(qlg_nrm2(ctx_->get(), DtypeOf<Elem>(), (int64_t)size_, dev_, &n), ctx_->get());
    return n;
  }
  double Normalize() {
    const double n = Norm();
    if (n != 0.0) {
      double inv[2] = {1.0 / n, 0.0};
      Check(qlg_scal(ctx_->get(), DtypeOf<Elem>(), (int64_t)size_, inv, dev_), ctx_->get());
      Check(qlg_ctx_synchronize(ctx_->get()), ctx_->get());
    }
    return n;
  }
  double QuasiNormalize() { return Normalize(); }

  // C descriptor (valid while this tensor is alive and unchanged)
  const qlg_bsten &Descriptor() const {
    d_nsect_.clear();
    d_qn_.clear();
    d_deg_.clear();
    d_dir_.clear();
    d_coords_.clear();
    const int qn_len = idx_.empty() || idx_[0].sectors().empty() ? 1 : (int)idx_[0].sectors()[0].qn.size();
    for (auto &ix : idx_) {
      d_nsect_.push_back((int32_t)ix.sectors().size());
      for (auto &s : ix.sectors()) {
        for (int c = 0; c < qn_len; ++c) d_qn_.push_back(s.qn[c]);
        d_deg_.push_back((int64_t)s.deg);
      }
      d_dir_.push_back((int8_t)ix.GetDir());
    }
    for (auto &c : blk_) d_coords_.insert(d_coords_.end(), c.begin(), c.end());
    desc_ = qlg_bsten{(int32_t)idx_.size(), qn_len, d_nsect_.data(), d_qn_.data(), d_deg_.data(), d_dir_.data(),
                      (int64_t)blk_.size(), d_coords_.data(), off_.data(), DtypeOf<Elem>(), dev_};
    return desc_;
  }

 private:
  void Alloc(size_t n) {
    Release();
    void *p = nullptr;
    Check(qlg_device_alloc(ctx_->get(), (int64_t)std::max<size_t>(n, 1) * sizeof(Elem), &p), ctx_->get());
    dev_ = static_cast<Elem *>(p);
    size_ = n;
  }
  void Release() {
    if (dev_ && ctx_) qlg_device_free(ctx_->get(), dev_);
    dev_ = nullptr;
    size_ = 0;
  }
  void Swap(BSTensor &o) {
    std::swap(idx_, o.idx_);
    std::swap(blk_, o.blk_);
    std::swap(off_, o.off_);
    std::swap(host_, o.host_);
    std::swap(dev_, o.dev_);
    std::swap(size_, o.size_);
    std::swap(ctx_, o.ctx_);
  }

  std::vector<Index> idx_;
  std::vector<Coords> blk_;
  std::vector<int64_t> off_;
  std::map<Coords, std::vector<Elem>> host_;
  Elem *dev_ = nullptr;
  size_t size_ = 0;
  Context *ctx_ = &Context::Default();
  mutable std::vector<int32_t> d_nsect_, d_qn_, d_coords_;
  mutable std::vector<int64_t> d_deg_;
  mutable std::vector<int8_t> d_dir_;
  mutable qlg_bsten desc_{};
};

// ---------------------------------------------------------------------------
// parameters and results (reference names)
// ---------------------------------------------------------------------------
struct LanczosParams {
  LanczosParams(double err, size_t max_iter) : error(err), max_iterations(max_iter) {}
  LanczosParams(double err) : LanczosParams(err, 200) {}
  LanczosParams(void) : LanczosParams(1.0E-7, 200) {}
  double error;
  size_t max_iterations;
};

template <typename TenT>
struct LanczosRes {
  size_t iters;
  double gs_eng;
  TenT *gs_vec;
};

template <typename TenT>
struct ExpmvRes {
  size_t iters;
  TenT *expmv;
};

// A planned effective Hamiltonian: sector pairing done once.
// DMRG-form term groups, include/qlmps/algorithm/dmrg/dmrg.h:16-45 (same
// names and nesting): group j pairs {L_i, W_l(i, j)} with {W_r(j, k), R_k}.
template <typename TenT>
using BlockSiteHamiltonianTerm = std::array<TenT *, 2>;
template <typename TenT>
using BlockSiteHamiltonianTermGroup = std::vector<BlockSiteHamiltonianTerm<TenT>>;
template <typename TenT>
using SiteBlockHamiltonianTerm = std::array<TenT *, 2>;
template <typename TenT>
using SiteBlockHamiltonianTermGroup = std::vector<SiteBlockHamiltonianTerm<TenT>>;
template <typename TenT>
using SuperBlockHamiltonianTerms =
    std::vector<std::pair<BlockSiteHamiltonianTermGroup<TenT>, SiteBlockHamiltonianTermGroup<TenT>>>;

template <typename Elem>
class EffectiveHamiltonian {
 public:
  using TenT = BSTensor<Elem>;
  EffectiveHamiltonian(const std::vector<TenT *> &eff_ham, const TenT &state, int precision = QLG_PREC_F64)
      : ctx_(&state.ctx()) {
    const qlg_ctx c = ctx_->get();
    if (eff_ham.size() == 4) {
      Check(qlg_plan_two_site(c, &eff_ham[0]->Descriptor(), &eff_ham[1]->Descriptor(), &eff_ham[2]->Descriptor(),
                              &eff_ham[3]->Descriptor(), &state.Descriptor(), precision, &plan_),
            c);
    } else if (eff_ham.size() == 3) {
      Check(qlg_plan_one_site(c, &eff_ham[0]->Descriptor(), &eff_ham[1]->Descriptor(), &eff_ham[2]->Descriptor(),
                              &state.Descriptor(), precision, &plan_),
            c);
    } else {
      throw std::runtime_error("effective Hamiltonian must be {L, Wl, Wr, R} or {L, W, R}");
    }
    ReadLayout(state);
  }
  // DMRG form: the term groups of SuperBlockHamiltonianTerms (see below).
  EffectiveHamiltonian(const SuperBlockHamiltonianTerms<TenT> &terms, const TenT &state,
                       int precision = QLG_PREC_F64)
      : ctx_(&state.ctx()) {
    const qlg_ctx c = ctx_->get();
    std::vector<int64_t> nl, nr;
    std::vector<qlg_term> lt, rt;
    for (const auto &g : terms) {
      nl.push_back((int64_t)g.first.size());
      nr.push_back((int64_t)g.second.size());
      for (const auto &t : g.first) lt.push_back(qlg_term{&t[0]->Descriptor(), &t[1]->Descriptor()});
      for (const auto &t : g.second) rt.push_back(qlg_term{&t[0]->Descriptor(), &t[1]->Descriptor()});
    }
    Check(qlg_plan_super_block(c, (int64_t)terms.size(), nl.data(), lt.data(), nr.data(), rt.data(),
                               &state.Descriptor(), precision, &plan_),
          c);
    ReadLayout(state);
  }

 private:
  void ReadLayout(const TenT &state) {
    const qlg_ctx c = ctx_->get();
    int64_t nb = 0, sz = 0;
    Check(qlg_plan_vector_layout(plan_, &nb, &sz, nullptr, nullptr), c);
    std::vector<int32_t> flat(nb * state.Rank());
    off_.resize(nb);
    Check(qlg_plan_vector_layout(plan_, nullptr, nullptr, flat.data(), off_.data()), c);
    for (int64_t b = 0; b < nb; ++b)
      blk_.emplace_back(flat.begin() + b * state.Rank(), flat.begin() + (b + 1) * state.Rank());
    size_ = (size_t)sz;
    idx_ = state.GetIndexes();
  }

 public:
  ~EffectiveHamiltonian() { qlg_plan_destroy(plan_); }
  EffectiveHamiltonian(const EffectiveHamiltonian &) = delete;
  EffectiveHamiltonian &operator=(const EffectiveHamiltonian &) = delete;

  // A new tensor in the plan's Krylov layout holding state's blocks.
  TenT *Adopt(const TenT &state) const {
    auto *t = new TenT(idx_, *ctx_);
    t->SetLayout(blk_, off_, size_);
    if (state.Blocks() == blk_ && state.Offsets() == off_) {
      Check(qlg_memcpy_d2d(ctx_->get(), t->DeviceData(), state.DeviceData(), size_ * sizeof(Elem)), ctx_->get());
      return t;
    }
    std::vector<Elem> zeros(size_, Elem(0));
    Check(qlg_memcpy_h2d(ctx_->get(), t->DeviceData(), zeros.data(), size_ * sizeof(Elem)), ctx_->get());
    std::map<std::vector<int32_t>, int64_t> pos;
    for (size_t b = 0; b < blk_.size(); ++b) pos[blk_[b]] = off_[b];
    for (size_t b = 0; b < state.Blocks().size(); ++b) {
      auto it = pos.find(state.Blocks()[b]);
      if (it == pos.end()) throw std::runtime_error("state block outside the Krylov layout");
      Check(qlg_memcpy_d2d(ctx_->get(), t->DeviceData() + it->second, state.DeviceData() + state.Offsets()[b],
                           state.BlockSize(state.Blocks()[b]) * sizeof(Elem)),
            ctx_->get());
    }
    return t;
  }
  TenT *Apply(const TenT &state) const {
    std::unique_ptr<TenT> in(Adopt(state));
    auto *out = new TenT(idx_, *ctx_);
    out->SetLayout(blk_, off_, size_);
    Check(qlg_matvec(plan_, in->DeviceData(), out->DeviceData()), ctx_->get());
    Check(qlg_ctx_synchronize(ctx_->get()), ctx_->get());
    return out;
  }
  qlg_plan plan() const { return plan_; }
  double AlgorithmicFlops() const {
    double f = 0;
    Check(qlg_plan_flops(plan_, &f, nullptr), ctx_->get());
    return f;
  }

 private:
  Context *ctx_;
  qlg_plan plan_ = nullptr;
  std::vector<std::vector<int32_t>> blk_;
  std::vector<int64_t> off_;
  size_t size_ = 0;
  std::vector<Index> idx_;
};

// ---------------------------------------------------------------------------
// matvec function-pointer targets (the reference's plug-in interface)
// ---------------------------------------------------------------------------
template <typename Elem>
BSTensor<Elem> *eff_ham_mul_two_site_state(const std::vector<BSTensor<Elem> *> &eff_ham, BSTensor<Elem> *state) {
  EffectiveHamiltonian<Elem> H(eff_ham, *state);
  return H.Apply(*state);
}

template <typename Elem>
BSTensor<Elem> *eff_ham_mul_single_site_state(const std::vector<BSTensor<Elem> *> &eff_ham,
                                              BSTensor<Elem> *state) {
  EffectiveHamiltonian<Elem> H(eff_ham, *state);
  return H.Apply(*state);
}

// ---------------------------------------------------------------------------
// solvers
// ---------------------------------------------------------------------------
// The whole Krylov loop runs in the library, so the function pointer must be
// this layer's own matvec for the operator count given: eff_ham_mul_two_site_state
// with {L, W_l, W_r, R} or eff_ham_mul_single_site_state with {L, W, R}.  Any
// other pointer (a caller's own matvec, such as the reference's noise
// expansion, two_site_update_noise_finite_vmps_impl.h:269-273) cannot be
// offloaded and is refused with std::invalid_argument; the initial state is
// then left to the caller.
template <typename TenT>
void CheckOffloadableMatvec(const std::vector<TenT *> &rpeff_ham,
                            TenT *(*eff_ham_mul_state)(const std::vector<TenT *> &, TenT *), const char *who) {
  using Elem = typename std::remove_const<
      typename std::remove_pointer<decltype(std::declval<TenT &>().DeviceData())>::type>::type;
  static_assert(std::is_same<TenT, BSTensor<Elem>>::value, "the solver runs on qlmps_amd::BSTensor");
  const bool two = eff_ham_mul_state == &eff_ham_mul_two_site_state<Elem> && rpeff_ham.size() == 4;
  const bool one = eff_ham_mul_state == &eff_ham_mul_single_site_state<Elem> && rpeff_ham.size() == 3;
  if (!two && !one)
    throw std::invalid_argument(std::string(who) +
                                ": the matvec must be eff_ham_mul_two_site_state (4 operators) or "
                                "eff_ham_mul_single_site_state (3 operators); other matvecs cannot run in the library");
}

template <typename TenT>
LanczosRes<TenT> LanczosSolver(const std::vector<TenT *> &rpeff_ham, TenT *pinit_state,
                               TenT *(*eff_ham_mul_state)(const std::vector<TenT *> &, TenT *),
                               const LanczosParams &params) {
  CheckOffloadableMatvec(rpeff_ham, eff_ham_mul_state, "LanczosSolver");
  using Elem = typename std::remove_pointer<decltype(pinit_state->DeviceData())>::type;
  EffectiveHamiltonian<typename std::remove_const<Elem>::type> H(rpeff_ham, *pinit_state);
  TenT *work = H.Adopt(*pinit_state);
  delete pinit_state;  // the solver consumes the initial state
  qlg_lanczos_params prm{params.error, (int64_t)params.max_iterations};
  qlg_lanczos_result res{};
  const int rc = qlg_lanczos_gs(H.plan(), work->DeviceData(), &prm, &res);
  if (rc != QLG_OK) {
    const qlg_ctx c = work->ctx().get();
    delete work;
    Check(rc, c);
  }
  return LanczosRes<TenT>{(size_t)res.iters, res.gs_eng, work};
}

// DMRG overload, include/qlmps/algorithm/dmrg/lanczos_dmrg_solver_impl.h:61-196:
// the same recurrence on the super-block operator; the initial state is
// consumed and gs_vec is new-ed for the caller.
template <typename TenT>
LanczosRes<TenT> LanczosSolver(SuperBlockHamiltonianTerms<TenT> &terms, TenT *pinit_state,
                               const LanczosParams &params) {
  using Elem = typename std::remove_pointer<decltype(pinit_state->DeviceData())>::type;
  EffectiveHamiltonian<typename std::remove_const<Elem>::type> H(terms, *pinit_state);
  TenT *work = H.Adopt(*pinit_state);
  delete pinit_state;
  qlg_lanczos_params prm{params.error, (int64_t)params.max_iterations};
  qlg_lanczos_result res{};
  const int rc = qlg_lanczos_gs(H.plan(), work->DeviceData(), &prm, &res);
  if (rc != QLG_OK) {
    const qlg_ctx c = work->ctx().get();
    delete work;
    Check(rc, c);
  }
  return LanczosRes<TenT>{(size_t)res.iters, res.gs_eng, work};
}

template <typename TenT>
ExpmvRes<TenT> LanczosExpmvSolver(const std::vector<TenT *> &rpeff_ham, TenT *pinit_state,
                                  TenT *(*eff_ham_mul_state)(const std::vector<TenT *> &, TenT *),
                                  const double step_length, const LanczosParams &params) {
  CheckOffloadableMatvec(rpeff_ham, eff_ham_mul_state, "LanczosExpmvSolver");
  using Elem = typename std::remove_pointer<decltype(pinit_state->DeviceData())>::type;
  EffectiveHamiltonian<typename std::remove_const<Elem>::type> H(rpeff_ham, *pinit_state);
  TenT *work = H.Adopt(*pinit_state);
  delete pinit_state;
  qlg_lanczos_params prm{params.error, (int64_t)params.max_iterations};
  int64_t iters = 0;
  const int rc = qlg_lanczos_expmv(H.plan(), work->DeviceData(), step_length, &prm, &iters);
  if (rc != QLG_OK) {
    const qlg_ctx c = work->ctx().get();
    delete work;
    Check(rc, c);
  }
  return ExpmvRes<TenT>{(size_t)iters, work};
}

// lanczos_vmps_solver_impl.h:306-355 (same signature; gs_vec new[]'d for 'V')
inline void TridiagGsSolver(const std::vector<double> &a, const std::vector<double> &b, const size_t n,
                            double &gs_eng, double *&gs_vec, const char jobz) {
  if (jobz == 'V') {
    gs_vec = new double[n];
    Check(qlg_tridiag_gs(a.data(), b.data(), (int64_t)n, 'V', &gs_eng, gs_vec));
  } else {
    Check(qlg_tridiag_gs(a.data(), b.data(), (int64_t)n, jobz, &gs_eng, nullptr));
  }
}

// lanczos_expmv_solver_impl.h:199-240
inline void TridiagExpme1Solver(const std::vector<double> &a, const std::vector<double> &b, const size_t n,
                                const double step_length, QLTEN_Complex *res) {
  static_assert(sizeof(QLTEN_Complex) == 2 * sizeof(double), "interleaved complex expected");
  Check(qlg_tridiag_expme1(a.data(), b.data(), (int64_t)n, step_length, reinterpret_cast<double *>(res)));
}

// lanczos_expmv_solver_impl.h:252-269 (v1 is left unchanged here)
template <typename ElemType>
inline double Distance(const ElemType *v1, const ElemType *v2, const size_t n) {
  double s = 0;
  for (size_t i = 0; i < n; ++i) s += std::norm(v2[i] - v1[i]);
  return std::sqrt(s);
}

}  // namespace qlmps_amd

#endif  // QLMPS_AMD_H
