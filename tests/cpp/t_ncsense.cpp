// t_ncsense.cpp -- drives the codeare-shaped C++ operators (include/codeare_amd/ncsense.hpp)
// the way codeare's CGSENSE module drives NCSENSE<cxfl> (src/modules/CGSENSE.cpp:132-178):
// construct from Params, KSpace, Weights, then *, / and ->*.
// Inputs are raw little-endian files written by tests/test_cpp_facade.py; outputs are
// written next to them for the test to compare with the CPU oracle.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "codeare_amd/ncsense.hpp"

template <class T>
struct MiniMatrix {   // column-major storage with the members the operators use
  explicit MiniMatrix(const std::vector<size_t>& dims) {
    size_t n = 1;
    for (size_t v : dims) n *= v;
    data.assign(n, T());
  }
  T* Ptr() { return data.data(); }
  const T* Ptr() const { return data.data(); }
  size_t Size() const { return data.size(); }
  std::vector<T> data;
};

using CMat = MiniMatrix<codeare_amd::cxfl>;
using RMat = MiniMatrix<float>;

template <class M>
static void load(const std::string& path, M& m) {
  std::ifstream f(path, std::ios::binary);
  if (!f) { std::fprintf(stderr, "missing %s\n", path.c_str()); std::exit(2); }
  f.read((char*)m.Ptr(), sizeof(m.data[0]) * m.Size());
}
template <class M>
static void save(const std::string& path, const M& m) {
  std::ofstream f(path, std::ios::binary);
  f.write((const char*)m.Ptr(), sizeof(m.data[0]) * m.Size());
}

// 3rd_dim_cart: Nz slices on the one trajectory (NCSENSE.hpp:81-99, NFFT.hpp:100-103)
static int run_slices(const std::string& dir, size_t N, size_t C, size_t nk, double m, double alpha, size_t Nz) {
  codeare_amd::Params p;
  p.set("nk", (double)nk);
  p.set("m", m);
  p.set("alpha", alpha);
  p.set("ftiter", 3);
  p.set("fteps", 1e-3);
  p.set("cgiter", 5);
  p.set("cgeps", 1e-7);
  p.set("lambda", 5e-2);
  p.set("3rd_dim_cart", 1);
  CMat S({N, N, Nz, C}), x({N, N, Nz}), y({nk, Nz, C});
  RMat k({2, nk}), w({nk});
  load(dir + "/S3.bin", S);
  load(dir + "/x3.bin", x);
  load(dir + "/y3.bin", y);
  load(dir + "/k.bin", k);
  load(dir + "/w.bin", w);
  codeare_amd::NCSENSE<CMat, RMat> ncs(p, S, {N, N, Nz, C});
  ncs.KSpace(k);
  ncs.Weights(w);
  save(dir + "/E3.bin", ncs * x);
  save(dir + "/EH3.bin", ncs / y);
  save(dir + "/solve3.bin", ncs->*y);
  codeare_amd::Params q;
  q.set("nk", (double)nk);
  q.set("m", m);
  q.set("alpha", alpha);
  q.set("3rd_dim_cart", 1);
  q.set_imsz({N, N, Nz});
  codeare_amd::NFFT<CMat, RMat> ft(q);
  ft.KSpace(k);
  ft.Weights(w);
  save(dir + "/trafo3.bin", ft * x);
  CMat y0({nk, Nz});
  for (size_t z = 0; z < Nz; ++z)
    for (size_t j = 0; j < nk; ++j) y0.data[j + nk * z] = y.data[j + nk * z];   // coil 0 of every slice
  save(dir + "/inv3.bin", ft.Adjoint(y0));
  return 0;
}

// GRASP-style frames: NCSENSE with dim4 x dim5 frames, per-frame trajectories
// (2, nk, dim4, dim5), shared weights, cgiter 0 and ftiter 1, so that ->* is the
// one-step reference E^H of every frame (GRASP.cpp:205-219 through
// CS_XSENSE.hpp:93-114; NCSENSE.hpp:141-151, :204-236, :254-288)
static int run_frames(const std::string& dir, size_t N, size_t C, size_t nk, double m, double alpha, size_t d4,
                      size_t d5) {
  codeare_amd::Params p;
  p.set("nk", (double)nk);
  p.set("m", m);
  p.set("alpha", alpha);
  p.set("ftiter", 1);
  p.set("fteps", 7e-4);
  p.set("cgiter", 0);
  p.set("dim4", (double)d4);
  p.set("dim5", (double)d5);
  CMat S({N, N, C}), x({N, N, d4, d5}), y({nk, C, d4, d5});
  RMat k({2, nk, d4, d5}), w({nk});
  load(dir + "/Sf.bin", S);
  load(dir + "/xf.bin", x);
  load(dir + "/yf.bin", y);
  load(dir + "/kf.bin", k);
  load(dir + "/wf.bin", w);
  codeare_amd::NCSENSE<CMat, RMat> ncs(p, S, {N, N, C});
  ncs.KSpace(k);
  ncs.Weights(w);
  CMat e = ncs * x;
  if (e.Size() != nk * C * d4 * d5) { std::fprintf(stderr, "frames: E has the wrong size\n"); return 4; }
  save(dir + "/Ef.bin", e);
  save(dir + "/EHf.bin", ncs / y);
  save(dir + "/solvef.bin", ncs->*y);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 7) { std::fprintf(stderr, "usage: t_ncsense dir N C nk m alpha [Nz [dim4 dim5]]\n"); return 2; }
  const std::string dir = argv[1];
  const size_t N = std::atoi(argv[2]), C = std::atoi(argv[3]), nk = std::atoi(argv[4]);
  codeare_amd::Params p;
  p.set("nk", (double)nk);
  p.set("m", std::atof(argv[5]));
  p.set("alpha", std::atof(argv[6]));
  p.set("ftiter", 3);
  p.set("fteps", 1e-3);
  p.set("cgiter", 5);
  p.set("cgeps", 1e-7);
  p.set("lambda", 5e-2);
  CMat S({N, N, C}), x({N, N}), y({nk, C});
  RMat k({2, nk}), w({nk});
  load(dir + "/S.bin", S);
  load(dir + "/x.bin", x);
  load(dir + "/y.bin", y);
  load(dir + "/k.bin", k);
  load(dir + "/w.bin", w);
  try {
    codeare_amd::NCSENSE<CMat, RMat> ncs(p, S, {N, N, C});
    ncs.KSpace(k);
    ncs.Weights(w);
    save(dir + "/E.bin", ncs * x);
    save(dir + "/EH.bin", ncs / y);
    save(dir + "/solve.bin", ncs->*y);
    // the single-coil NFFT<cxfl>: Trafo and the iterative Adjoint
    codeare_amd::Params q;
    q.set("nk", (double)nk);
    q.set("m", std::atof(argv[5]));
    q.set("alpha", std::atof(argv[6]));
    q.set_imsz({N, N});
    codeare_amd::NFFT<CMat, RMat> ft(q);
    ft.KSpace(k);
    ft.Weights(w);
    save(dir + "/trafo.bin", ft * x);
    CMat y0({nk});
    for (size_t j = 0; j < nk; ++j) y0.data[j] = y.data[j];
    save(dir + "/inv.bin", ft.Adjoint(y0));
    // errors cross the boundary as exceptions carrying the library message
    bool threw = false;
    try {
      RMat bad({2, nk});
      for (auto& v : bad.data) v = 0.75f;   // outside [-0.5, 0.5)
      ft.KSpace(bad);
    } catch (const std::runtime_error& e) {
      threw = std::string(e.what()).find("outside") != std::string::npos;
    }
    if (!threw) { std::fprintf(stderr, "bad trajectory not rejected\n"); return 3; }
    if (argc >= 8 && std::atoi(argv[7]) > 0)
      run_slices(dir, N, C, nk, std::atof(argv[5]), std::atof(argv[6]), (size_t)std::atoi(argv[7]));
    if (argc >= 10 && std::atoi(argv[8]) > 0 && std::atoi(argv[9]) > 0) {
      const int rc = run_frames(dir, N, C, nk, std::atof(argv[5]), std::atof(argv[6]), (size_t)std::atoi(argv[8]),
                                (size_t)std::atoi(argv[9]));
      if (rc != 0) return rc;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
