// The reference's stencil test scenarios (test/shared-stencil.hpp:120-239)
// written against the header-only GPU classes: a 4 x 9 x 7 lookup stencil
// with the diagonal-lines materials (ids (y + z) % 4, coefficients 0..3,
// test/aux/materials.hpp:217-233).
//   known solution 2: PCG-AMG and BiCGStab-AMG reach residual <= 1e-13 and
//                     |x - 2|_inf <= 1e-13 within 32 iterations
//   the reference's solver suite (test/aux/iterative.hpp:67-181) through
//                     run_iterative_solver with its cg and gs iteration
//                     callbacks (test/aux/iterative.hpp:25-65): cg, bicgstab,
//                     sor, ssor1, sor5, ssor5, pcg_ssor1, pbicgstab_ssor1,
//                     pcg_amg, pbicgstab_amg for both SOR variants, against
//                     the bounds of test/shared-stencil.hpp:157-168
//   Dirichlet rhs:    PCG-AMG residual |A x - b|_inf <= 2e-14
//   CSR copy:         CSR SpMV equals the matrix-free SpMV
//   adjoint:          outlet flux of a constant-coefficient values stencil
//                     equals n_y n_z c dp / n_x and dJ/dlog c_k = J_col^2 / c_k
//   effective diffusivity: the compressed 4 x 7 x 6 system of
//                     test/shared-csr.cpp:193-252 gives the mean coefficient
//   files (argv[1]: folder of UFZ_CT_02_LHS.smf / UFZ_CT_02_RHS.dvf): read,
//                     written back byte for byte (test/tensor-io.cpp:58-100),
//                     solved with AMG-CG
// Prints one line per check; the exit status is the number of failed checks.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "lineal-amd/lineal-amd.hpp"

namespace lg = lineal::gpu;

static int failures = 0;

static void expect(bool ok, const char* what, double value) {
  std::printf("%s %s (%.3e)\n", ok ? "ok  " : "FAIL", what, value);
  if (!ok) ++failures;
}

static std::vector<char> file_bytes(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  return {std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>()};
}

static double inf_norm_diff(const std::vector<double>& a, double v) {
  double m = 0;
  for (double x : a) m = std::fmax(m, std::fabs(x - v));
  return m;
}

// test/aux/iterative.hpp:25-45: iterate until iter > iter_num or rho == 0 and
// return the residual norm of the system instance
template <typename TSolver>
static double cg(const TSolver& solver, const lg::GpuMatrixBase& lhs, lg::GpuDenseVector<double>& sol,
                 const lg::GpuDenseVector<double>& rhs, std::size_t iter_num) {
  return lg::run_iterative_solver(solver, lhs, sol, rhs, [&](std::size_t iter, auto& inst) {
    const double res = inst.residual_norm();
    const double rho = inst.rho();
    if (!std::isfinite(res)) throw std::runtime_error("non-finite residual norm");
    return (iter > iter_num || rho == 0) ? std::optional<double>{res} : std::nullopt;
  });
}

// test/aux/iterative.hpp:47-65: the Euclidean norm of lhs * sol - rhs after
// every application of a stationary solver
template <typename TSolver>
static double gs(const TSolver& solver, const lg::GpuMatrixBase& lhs, lg::GpuDenseVector<double>& sol,
                 const lg::GpuDenseVector<double>& rhs, std::size_t iter_num) {
  lg::GpuDenseVector<double> r(sol.size(), sol.context());
  return lg::run_iterative_solver(solver, lhs, sol, rhs, [&](std::size_t iter, auto&) {
    lhs.residual(rhs, sol, r);
    double sq = 0;
    for (double v : r.to_host()) sq += v * v;
    return iter > iter_num ? std::optional<double>{std::sqrt(sq)} : std::nullopt;
  });
}

// test::suite (test/aux/iterative.hpp:67-181) with the bounds of
// test/shared-stencil.hpp:157-168; returns the number of failed solvers
template <lg::SorVariant tVariant>
static int suite(const lg::GpuStencilLookupMatrix& lhs, lg::GpuDenseVector<double>& sol,
                 const lg::GpuDenseVector<double>& rhs) {
  const std::vector<double> bounds{1e-12, 1e-12, 1e-8, 1e-13, 1e-13, 1e-13, 1e-13, 1e-12, 1e-13, 1e-13};
  const char* names[] = {"cg",    "bicgstab",  "sor",        "ssor1",   "sor5",
                         "ssor5", "pcg_ssor1", "pbicgstab_ssor1", "pcg_amg", "pbicgstab_amg"};
  int failed = 0;
  std::size_t k = 0;
  auto op = [&](double res) {
    const double err = inf_norm_diff(sol.to_host(), 2.0);
    const bool ok = res <= bounds[k] && err <= bounds[k];
    if (!ok) {
      std::printf("FAIL   %s: residual %.3e, error %.3e, bound %.1e\n", names[k], res, err, bounds[k]);
      ++failed;
    }
    ++k;
  };
  auto zero = [&] { sol.fill(0.0); };
  using Ssor = lg::GpuSsorSolver<double, tVariant>;
  const Ssor ssor1(1.0, 1);
  const lg::GpuMultiGridSolver<double> amg({}, tVariant);
  zero();
  op(cg(lg::GpuConjugateGradients<Ssor>(8), lhs, sol, rhs, 128));
  zero();
  op(cg(lg::GpuBiCgStab<Ssor>(8), lhs, sol, rhs, 128));
  zero();
  op(gs(lg::GpuSorSolver<double, lg::SorDirection::FORWARD, tVariant>(1.0), lhs, sol, rhs, 128));
  zero();
  op(gs(ssor1, lhs, sol, rhs, 128));
  zero();
  op(gs(lg::GpuSorSolver<double, lg::SorDirection::FORWARD, tVariant>(1.0, 5), lhs, sol, rhs, 128));
  zero();
  op(gs(Ssor(1.0, 5), lhs, sol, rhs, 128));
  zero();
  op(cg(lg::GpuConjugateGradients<Ssor>(8, &ssor1), lhs, sol, rhs, 128));
  zero();
  op(cg(lg::GpuBiCgStab<Ssor>(8, &ssor1), lhs, sol, rhs, 128));
  zero();
  op(cg(lg::GpuConjugateGradients<lg::GpuMultiGridSolver<double>>(8, &amg), lhs, sol, rhs, 32));
  zero();
  op(cg(lg::GpuBiCgStab<lg::GpuMultiGridSolver<double>>(8, &amg), lhs, sol, rhs, 32));
  return failed;
}

int main(int argc, char** argv) {
  try {
    lg::GpuContext ctx(0);
    const std::array<int64_t, 3> dims{4, 9, 7};
    lg::BoxInfo box(dims);
    std::vector<std::uint8_t> mat(size_t(box.cell_num()));
    for (int64_t x = 0; x < dims[0]; ++x)
      for (int64_t y = 0; y < dims[1]; ++y)
        for (int64_t z = 0; z < dims[2]; ++z)
          mat[size_t((x * dims[1] + y) * dims[2] + z)] = std::uint8_t((y + z) % 4);
    lg::GpuStencilLookupMatrix A(box, mat, {0.0, 1.0, 2.0, 3.0}, ctx);
    const size_t n = A.row_num();
    expect(A.is_symmetric() && n == 252, "lookup stencil rows", double(n));

    // known solution 2 (test/shared-stencil.hpp:154-190)
    lg::GpuDenseVector<double> two(n, 2.0, ctx);
    const auto b = A * two;
    lg::GpuMultiGridSolver<double> amg;
    for (int which = 0; which < 2; ++which) {
      lg::GpuDenseVector<double> x(n, 0.0, ctx);
      lg::SolveReport r;
      if (which == 0) {
        lg::GpuConjugateGradients<> cg(8, &amg);
        r = lg::run_iterative_solver(cg, A, x, b, {1e-15, 32});
      } else {
        lg::GpuBiCgStab<> bicg(8, &amg);
        r = lg::run_iterative_solver(bicg, A, x, b, {1e-15, 32});
      }
      const double err = inf_norm_diff(x.to_host(), 2.0);
      expect(r.residual_norm() <= 1e-13 && err <= 1e-13,
             which == 0 ? "pcg_amg known solution 2" : "bicgstab_amg known solution 2", err);
    }

    // the reference's solver suite through iteration callbacks
    {
      lg::GpuDenseVector<double> sol(n, 0.0, ctx);
      const int fr = suite<lg::SorVariant::REGULAR>(A, sol, b);
      expect(fr == 0, "reference suite through callbacks, regular SOR", double(fr));
      const int fu = suite<lg::SorVariant::ULTRA>(A, sol, b);
      expect(fu == 0, "reference suite through callbacks, ultra SOR", double(fu));
    }

    // Dirichlet rhs (test/shared-stencil.hpp:192-239)
    {
      const auto rhs = A.rhs();
      lg::GpuDenseVector<double> x(n, 0.0, ctx);
      lg::GpuConjugateGradients<> cg(8, &amg);
      const auto r = lg::run_iterative_solver(cg, A, x, rhs, {1e-15, 64});
      lg::GpuDenseVector<double> res(n, ctx);
      A.residual(rhs, x, res);
      const double rinf = inf_norm_diff(res.to_host(), 0.0);
      expect(rinf <= 2e-14, "pcg_amg dirichlet residual", rinf);
      // ssor1 (129 applications, 1.1e-14) and pcg_ssor1 (1e-13),
      // test/aux/iterative.hpp:122-153 with test/shared-stencil.hpp:198-224
      lg::GpuSsorSolver<double> ssor(1.0);
      const auto si = ssor.instantiate(A);
      lg::GpuDenseVector<double> y(n, 0.0, ctx);
      for (int k = 0; k < 129; ++k) si.apply(&y, y, rhs);
      A.residual(rhs, y, res);
      const double sinf = inf_norm_diff(res.to_host(), 0.0);
      lg::GpuDenseVector<double> z(n, 0.0, ctx);
      lg::GpuConjugateGradients<lg::GpuSsorSolver<double>> pcg(8, &ssor);
      const auto rp = lg::run_iterative_solver(pcg, A, z, rhs, {1e-300, 129});
      A.residual(rhs, z, res);
      const double pinf = inf_norm_diff(res.to_host(), 0.0);
      expect(sinf <= 1.1e-14 && pinf <= 1e-13 && rp.residual_norm() <= 1e-13,
             "ssor1 and pcg_ssor1 dirichlet residual", std::fmax(sinf, pinf));
      (void)r;
    }

    // the SOR sweeps' coefficient cache changes no bit
    {
      const auto rhs = A.rhs();
      lg::GpuSsorSolver<double> ssor(1.0);
      const auto si = ssor.instantiate(A);
      lg::GpuDenseVector<double> y0(n, 0.0, ctx), y1(n, 0.0, ctx);
      for (int k = 0; k < 5; ++k) si.apply(&y0, y0, rhs);
      A.set_sor_cache(false);
      for (int k = 0; k < 5; ++k) si.apply(&y1, y1, rhs);
      A.set_sor_cache(true);
      const auto h0 = y0.to_host(), h1 = y1.to_host();
      double diff = 0.0;
      bool same = true;
      for (size_t i = 0; i < n; ++i)
        if (h0[i] != h1[i]) {
          same = false;
          diff = std::fmax(diff, std::fabs(h0[i] - h1[i]));
        }
      expect(same, "ssor with and without the sor coefficient cache", diff);
    }

    // CSR copy of the stencil
    {
      lg::GpuCsrMatrix<double> C(A);
      std::vector<double> xv(n);
      for (size_t i = 0; i < n; ++i) xv[i] = std::sin(double(i));
      lg::GpuDenseVector<double> x(std::span<const double>(xv), ctx);
      const auto y1 = (A * x).to_host();
      const auto y2 = (C * x).to_host();
      double d = 0;
      for (size_t i = 0; i < n; ++i) d = std::fmax(d, std::fabs(y1[i] - y2[i]));
      expect(d <= 1e-14, "csr spmv equals matrix-free spmv", d);
    }

    // stationary solvers: one forward SOR sweep on A x = A 2 from zero
    {
      lg::GpuSorSolver<double> sor(1.0);
      auto inst = sor.instantiate(A);
      lg::GpuDenseVector<double> x(n, ctx);
      inst.apply(nullptr, x, b);
      lg::GpuJacobiSolver<double> jac(2.0 / 3.0);
      lg::GpuDenseVector<double> y(n, ctx);
      jac.instantiate(A).apply(&two, y, b);
      const double e = inf_norm_diff(y.to_host(), 2.0);
      expect(e <= 1e-14, "jacobi keeps the exact solution", e);
    }

    // adjoint on a constant-coefficient values stencil: J = n_y n_z c dp / n_x
    {
      const std::array<int64_t, 3> d2{16, 6, 5};
      lg::BoxInfo b2(d2);
      std::vector<double> c(size_t(b2.cell_num()), 1.5);
      lg::GpuStencilValueMatrix V(b2, c, ctx);
      lg::GpuMultiGridSolver<double> m2;
      auto mi = m2.instantiate(V);
      const size_t n2 = V.row_num();
      lg::GpuDenseVector<double> p(n2, 0.0, ctx), lam(n2, 0.0, ctx), g(n2, ctx);
      const auto r = lg::outlet_flux_gradient(V, &mi, p, lam, g, {1e-13, 200});
      const double j = double(d2[1] * d2[2]) * 1.5 / double(d2[0]);
      expect(r.converged() && std::fabs(r.flux - j) <= 1e-11 * j, "outlet flux closed form",
             std::fabs(r.flux - j));
      // column closed form: dJ/dlog c_k = J_col^2 / c_k with J_col = c / n_x
      const double jc = 1.5 / double(d2[0]);
      const double e = inf_norm_diff(g.to_host(), jc * jc / 1.5);
      expect(e <= 1e-11, "outlet flux gradient closed form", e);
    }
    // effective diffusivity of the compressed system (test/shared-csr.cpp:193-252):
    // 4 x 7 x 6 diagonal lines, ZeroToZero, p_start = 1, p_end = 0; the mean
    // coefficient over a cross-section (test/aux/materials.hpp:235-248)
    {
      const std::array<int64_t, 3> d3{4, 7, 6};
      lg::BoxInfo b3(d3);
      b3.zero_to_one = false;
      std::vector<std::uint8_t> m3(size_t(b3.cell_num()));
      double mean = 0.0;
      for (int64_t x = 0; x < d3[0]; ++x)
        for (int64_t y = 0; y < d3[1]; ++y)
          for (int64_t z = 0; z < d3[2]; ++z) {
            const int id = int((y + z) % 4);
            m3[size_t((x * d3[1] + y) * d3[2] + z)] = std::uint8_t(id);
            if (x == 0) mean += double(id);
          }
      mean /= double(d3[1] * d3[2]);
      lg::GpuStencilLookupMatrix L(b3, m3, {0.0, 1.0, 2.0, 3.0}, ctx);
      auto sys = lg::csr_from_adjacent_stencil<double>(L, true);
      const size_t nc = sys.lhs.row_num();
      lg::GpuDenseVector<double> x(nc, 0.0, ctx);
      lg::GpuConjugateGradients<> cg(8, nullptr);
      const auto r = lg::run_iterative_solver(cg, sys.lhs, x, sys.rhs, {1e-14, 400});
      lg::GpuEffDiffCalc calc(L);
      const double eff = calc(L, sys, x);
      expect(r.converged() && nc < L.row_num() && std::fabs(eff - mean) <= 1e-10,
             "effective diffusivity of the compressed system", std::fabs(eff - mean));
    }
    if (argc > 1) {
      const std::string lhs = std::string(argv[1]) + "/UFZ_CT_02_LHS.smf";
      const std::string rhs = std::string(argv[1]) + "/UFZ_CT_02_RHS.dvf";
      const std::string tmp = (std::filesystem::temp_directory_path() /
                               ("lineal_amd_io_" + std::to_string(getpid())))
                                  .string();
      lg::SmfReader<double> reader(lhs);
      const auto U = reader.read(ctx, /*symmetric=*/true);
      lg::SmfWriter(tmp + ".smf").write(U);
      const auto b = lg::DvfReader<double>(rhs).read(ctx);
      lg::DvfWriter(tmp + ".dvf").write(b);
      const bool same = file_bytes(lhs) == file_bytes(tmp + ".smf") &&
                        file_bytes(rhs) == file_bytes(tmp + ".dvf");
      std::remove((tmp + ".smf").c_str());
      std::remove((tmp + ".dvf").c_str());
      lg::GpuDenseVector<double> x(U.row_num(), 0.0, ctx);
      lg::GpuMultiGridSolver<double> m3;
      lg::GpuConjugateGradients<> cg(8, &m3);
      const auto r = lg::run_iterative_solver(cg, U, x, b, {1e-10, 200});
      const double rel = r.residual_norms.back() / r.residual_norms.front();
      expect(same && reader.nnz_num() == 3367 && r.converged() && rel <= 1e-10,
             "SMF/DVF round trip and AMG-CG solve of UFZ_CT_02", rel);
    }
  } catch (const lg::GpuError& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 100;
  }
  return failures;
}
