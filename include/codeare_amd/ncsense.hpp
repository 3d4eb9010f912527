// ncsense.hpp -- header-only C++ operators over the C ABI, shaped like codeare's
// NFFT<cxfl> (src/matrix/ft/NFFT.hpp) and NCSENSE<cxfl> (src/matrix/ft/NCSENSE.hpp).
//
// The matrix type is a template parameter: anything column-major with
//   Matrix(const std::vector<size_t>& dims), T* Ptr(), const T* Ptr() const,
//   size_t Size() const
// works, which includes codeare's Matrix<cxfl> / Matrix<float>.  Parameters use
// codeare's keys and defaults (NCSENSE.hpp:69-184, NFFT.hpp:66-191):
//   NFFT:    nk, imsz, m (1), alpha (1.0), epsilon (7e-4), maxit (3),
//            3rd_dim_cart (0: with 1, imsz (N0, N1, Nz) is Nz 2D slices on one
//            trajectory, NFFT.hpp:84-103)
//   NCSENSE: nk, ftiter (3), fteps (1e-3), cgiter (0), cgeps (1e-6),
//            lambda (1e-6), m (1), alpha (1.0), dim4 (1), dim5 (1),
//            3rd_dim_cart (0: with 1, maps (N0, N1, Nz, C) and data (nk, Nz, C),
//            NCSENSE.hpp:81-99)
// Frames (dim4 x dim5 of them, frame l + dim4*n for index (l, n),
// NCSENSE.hpp:141-151, :222-233) and slices share one CG, as codeare's CGLS
// runs on the whole image matrix (cgs_solve_opts.joint).
// Errors from the library are thrown as std::runtime_error carrying the
// library's message (codeare asserts or throws NCSENSE_KSPACE_DIMENSIONS).
#pragma once

#include <complex>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../cgsense.h"

namespace codeare_amd {

using cxfl = std::complex<float>;

// numeric parameter set (the numeric part of codeare's Params)
class Params {
 public:
  void set(const std::string& k, double v) { v_[k] = v; }
  bool exists(const std::string& k) const { return v_.count(k) != 0; }
  double get(const std::string& k, double dflt) const {
    auto it = v_.find(k);
    return it == v_.end() ? dflt : it->second;
  }
  void set_imsz(const std::vector<size_t>& n) { imsz_ = n; }
  const std::vector<size_t>& imsz() const { return imsz_; }

 private:
  std::map<std::string, double> v_;
  std::vector<size_t> imsz_;
};

inline void check(int rc) {
  if (rc != CGS_OK) throw std::runtime_error(std::string(cgs_status_string(rc)) + ": " + cgs_last_error());
}

// NFFT<cxfl>: one plan, one coil.
template <class Matrix, class RealMatrix>
class NFFT {
 public:
  explicit NFFT(const Params& p) {
    if (!p.exists("nk") || p.imsz().empty()) throw std::runtime_error("NFFT: nk and imsz must be given");
    nk_ = (int64_t)p.get("nk", 0);
    for (size_t v : p.imsz()) N_.push_back((int)v);
    maxit_ = (int)p.get("maxit", 3);
    eps_ = p.get("epsilon", 7.0e-4);
    if (p.get("3rd_dim_cart", 0) != 0 && N_.size() == 3) {   // NFFT.hpp:100-103
      ncart_ = N_.back();
      N_.pop_back();
    }
    // the slices of a Cartesian third axis are the plan's image batch
    check(cgs_plan_create(&plan_, (int)N_.size(), N_.data(), nk_, p.get("alpha", 1.0), (int)p.get("m", 1), ncart_,
                          1, nullptr, 0));
  }
  ~NFFT() { cgs_plan_destroy(plan_); }
  NFFT(const NFFT&) = delete;
  NFFT& operator=(const NFFT&) = delete;

  void KSpace(const RealMatrix& k) {
    check(cgs_set_trajectory(plan_, k.Ptr()));
    if (!b0_.empty()) check(cgs_set_b0(plan_, b0_.data(), t_.data()));
  }
  void Weights(const RealMatrix& w) { check(cgs_set_weights(plan_, w.Ptr())); }
  // codeare's "b0" (N0, N1) and "timing" (nk) Params (NFFT.hpp:117-173): the
  // off-resonance model of Trafo and Adjoint, applied once KSpace is set
  void B0(const RealMatrix& b0, const RealMatrix& timing) {
    b0_.assign(b0.Ptr(), b0.Ptr() + b0.Size());
    t_.assign(timing.Ptr(), timing.Ptr() + timing.Size());
    if (b0_.size() != (size_t)N_[0] * (size_t)N_[1] || t_.size() != (size_t)nk_) {
      b0_.clear();
      t_.clear();
      throw std::runtime_error("NFFT: b0 must fit the image and timing the trajectory");
    }
    check(cgs_set_b0(plan_, b0_.data(), t_.data()));
  }

  Matrix Trafo(const Matrix& img) const {            // NFFT.hpp:305-361 (per slice)
    Matrix out(ncart_ > 1 ? std::vector<size_t>{(size_t)nk_, (size_t)ncart_} : std::vector<size_t>{(size_t)nk_});
    check(cgs_nfft_forward(plan_, 0, img.Ptr(), out.Ptr(), ncart_));
    return out;
  }
  Matrix operator*(const Matrix& img) const { return Trafo(img); }
  Matrix Adjoint(const Matrix& f) const {             // NFFT.hpp:370-436 (iterative inverse, per slice)
    Matrix out(image_dims());
    check(cgs_nfft_inverse_ref(plan_, 0, f.Ptr(), out.Ptr(), ncart_, maxit_, eps_));
    return out;
  }
  Matrix AdjointTrue(const Matrix& f) const {         // nfft_adjoint, NFFTTraits.hpp:201-204
    Matrix out(image_dims());
    check(cgs_nfft_adjoint(plan_, 0, f.Ptr(), out.Ptr(), ncart_));
    return out;
  }

 private:
  std::vector<size_t> image_dims() const {
    std::vector<size_t> d(N_.begin(), N_.end());
    if (ncart_ > 1) d.push_back((size_t)ncart_);
    return d;
  }
  cgs_plan plan_ = nullptr;
  std::vector<int> N_;
  std::vector<float> b0_, t_;
  int ncart_ = 1;
  int64_t nk_ = 0;
  int maxit_ = 3;
  double eps_ = 7.0e-4;
};

// NCSENSE<cxfl>: coil encoding, reference E^H and the CGLS solve.
template <class Matrix, class RealMatrix>
class NCSENSE {
 public:
  // sens: (N0, N1[, N2], C) complex maps, as Params["sensitivities"] in codeare
  NCSENSE(const Params& p, const Matrix& sens, const std::vector<size_t>& sens_dims) {
    if (sens_dims.size() < 3) throw std::runtime_error("NCSENSE: maps must be (N..., C)");
    cart_ = p.get("3rd_dim_cart", 0) != 0;
    if (cart_ && sens_dims.size() != 4) throw std::runtime_error("NCSENSE: 3rd_dim_cart needs (N0, N1, Nz, C) maps");
    d_ = (int)sens_dims.size() - (cart_ ? 2 : 1);
    C_ = (int)sens_dims.back();
    for (int a = 0; a < d_; ++a) N_.push_back((int)sens_dims[a]);
    nk_ = (int64_t)p.get("nk", 1);
    // frames of the plan: dim4 x dim5 frames (m_nmany, NCSENSE.hpp:141-151) or
    // the Cartesian slices, which share the trajectory, the weights and one CG
    dim4_ = cart_ ? 1 : (int)p.get("dim4", 1);
    dim5_ = cart_ ? 1 : (int)p.get("dim5", 1);
    if (dim4_ < 1 || dim5_ < 1) throw std::runtime_error("NCSENSE: dim4 and dim5 must be >= 1");
    nmany_ = cart_ ? (int)sens_dims[2] : dim4_ * dim5_;
    ftiter_ = (int)p.get("ftiter", 3);
    fteps_ = p.get("fteps", 1.0e-3);
    cgiter_ = (int)p.get("cgiter", 0);
    cgeps_ = p.get("cgeps", 1.0e-6);
    lambda_ = p.get("lambda", 1.0e-6);
    std::vector<int64_t> off;
    for (int f = 0; f <= nmany_; ++f) off.push_back(nk_ * f);
    check(cgs_plan_create(&plan_, d_, N_.data(), nk_ * nmany_, p.get("alpha", 1.0), (int)p.get("m", 1), C_,
                          nmany_, nmany_ > 1 ? off.data() : nullptr, cart_ ? CGS_MAPS_PER_FRAME : 0u));
    check(cgs_set_sensitivities(plan_, sens.Ptr()));
  }
  ~NCSENSE() { cgs_plan_destroy(plan_); }
  NCSENSE(const NCSENSE&) = delete;
  NCSENSE& operator=(const NCSENSE&) = delete;

  void Sensitivities(const Matrix& sm) { check(cgs_set_sensitivities(plan_, sm.Ptr())); }
  // k: (d, nk) shared by every frame/slice, or (d, nk, dim4[, dim5]) per frame
  // (NCSENSE.hpp:204-236); w: (nk) shared by every frame (NCSENSE.hpp:244-248)
  void KSpace(const RealMatrix& k) {
    check(cgs_set_trajectory(plan_, per_frame(k.Ptr(), k.Size(), (size_t)d_ * nk_).data()));
  }
  void Weights(const RealMatrix& w) { check(cgs_set_weights(plan_, per_frame(w.Ptr(), w.Size(), (size_t)nk_).data())); }

  Matrix Trafo(const Matrix& x) const {               // NCSENSE.hpp:310-348
    size_t nx = (size_t)nmany_;
    for (int v : N_) nx *= (size_t)v;
    if ((size_t)x.Size() != nx) throw std::runtime_error("NCSENSE: image size does not match N x frames");
    Matrix y(data_dims());
    if (!cart_) {
      check(cgs_E(plan_, x.Ptr(), y.Ptr()));
    } else {
      std::vector<cxfl> t((size_t)nk_ * C_ * nmany_);
      check(cgs_E(plan_, x.Ptr(), t.data()));
      slices_to_codeare(t.data(), y.Ptr());
    }
    return y;
  }
  Matrix operator*(const Matrix& x) const { return Trafo(x); }
  Matrix operator/(const Matrix& y) const {           // NCSENSE.hpp:251-301
    Matrix x(image_dims());
    check(cgs_EH(plan_, plan_data(y).data(), x.Ptr(), CGS_EH_REF, ftiter_, fteps_));
    return x;
  }
  Matrix Adjoint(const Matrix& y) const {             // NCSENSE.hpp:408-414 -> CGLS.hpp:30-59
    cgs_solve_opts o{};
    o.maxit = cgiter_;
    o.tol = cgeps_;
    o.tol_kind = CGS_TOL_SQUARED_RATIO;
    o.lambda = lambda_;
    o.eh_mode = CGS_EH_REF;
    o.ftiter = ftiter_;
    o.fteps = fteps_;
    o.normal_op = CGS_NORMAL_NUFFT;
    o.graph_iters = 0;
    o.joint = 1;
    Matrix x(image_dims());
    const int rc = cgs_solve(plan_, plan_data(y).data(), nullptr, x.Ptr(), &o, nullptr);
    if (rc != CGS_ERR_NAN) check(rc);   // codeare breaks out of CGLS on NaN and returns x
    return x;
  }
  Matrix operator->*(const Matrix& y) const { return Adjoint(y); }

  cgs_plan plan() const { return plan_; }

 private:
  // (N..., Nz) for slices; (N..., dim4[, dim5]) for frames (m_bwd_out, NCSENSE.hpp:172-179)
  std::vector<size_t> image_dims() const {
    std::vector<size_t> d(N_.begin(), N_.end());
    if (cart_) {
      d.push_back((size_t)nmany_);
    } else if (nmany_ > 1) {
      d.push_back((size_t)dim4_);
      if (dim5_ > 1) d.push_back((size_t)dim5_);
    }
    return d;
  }
  // (nk, Nz, C) for slices; (nk, C[, dim4[, dim5]]) for frames (m_fwd_out, NCSENSE.hpp:171)
  std::vector<size_t> data_dims() const {
    if (cart_) return std::vector<size_t>{(size_t)nk_, (size_t)nmany_, (size_t)C_};
    std::vector<size_t> d{(size_t)nk_, (size_t)C_};
    if (dim4_ > 1) d.push_back((size_t)dim4_);
    if (dim5_ > 1) d.push_back((size_t)dim5_);
    return d;
  }
  // one block per frame when the caller gives one shared block
  template <class T>
  std::vector<T> per_frame(const T* v, size_t n, size_t block) const {
    if (n != block && n != block * nmany_) throw std::runtime_error("NCSENSE: k-space / weights size mismatch");
    std::vector<T> out(block * nmany_);
    for (int f = 0; f < nmany_; ++f)
      for (size_t i = 0; i < block; ++i) out[f * block + i] = v[n == block ? i : f * block + i];
    return out;
  }
  // codeare's slice data (nk, Nz, C) <-> the plan's frame blocks (nk, C) x Nz
  void slices_to_codeare(const cxfl* t, cxfl* y) const {
    for (int z = 0; z < nmany_; ++z)
      for (int c = 0; c < C_; ++c)
        for (int64_t j = 0; j < nk_; ++j) y[j + nk_ * (z + (int64_t)nmany_ * c)] = t[j + nk_ * (c + (int64_t)C_ * z)];
  }
  std::vector<cxfl> plan_data(const Matrix& y) const {
    if ((size_t)y.Size() != (size_t)nk_ * C_ * nmany_)
      throw std::runtime_error("NCSENSE: data size does not match nk x coils x frames");
    const cxfl* s = reinterpret_cast<const cxfl*>(y.Ptr());
    std::vector<cxfl> t(s, s + (size_t)nk_ * C_ * nmany_);
    if (cart_)
      for (int z = 0; z < nmany_; ++z)
        for (int c = 0; c < C_; ++c)
          for (int64_t j = 0; j < nk_; ++j) t[j + nk_ * (c + (int64_t)C_ * z)] = s[j + nk_ * (z + (int64_t)nmany_ * c)];
    return t;
  }
  cgs_plan plan_ = nullptr;
  bool cart_ = false;
  int d_ = 2, C_ = 1, nmany_ = 1, dim4_ = 1, dim5_ = 1;
  std::vector<int> N_;
  int64_t nk_ = 1;
  int ftiter_ = 3, cgiter_ = 0;
  double fteps_ = 1.0e-3, cgeps_ = 1.0e-6, lambda_ = 1.0e-6;
};

}  // namespace codeare_amd
