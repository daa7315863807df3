// Oracle Krylov solver -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Restates the solver semantics of src/linear/backends/PetscLinearSolver.cpp:186-220:
// KSP "cg" with x0 = 0 (KSPSetInitialGuessNonzero is never called), tolerances re-applied on
// every solve, failure (reason < 0, here: max_it exhausted or breakdown) raised as an error.
// The preconditioner is Jacobi for both operators; the reference's GW default "sor" is a
// sequential Gauss-Seidel sweep that converges to the same discrete solution.
#include <algorithm>
#include <stdexcept>

#include "oracle.hpp"

namespace frehg_oracle {

real det_sum(const std::vector<real>& v) {
  const long long n = (long long)v.size();
  const long long B = 4096;
  const long long nb = (n + B - 1) / B;
  std::vector<real> part((size_t)nb, 0.0);
#pragma omp parallel for schedule(static) if(nb >= 8)
  for (long long b = 0; b < nb; ++b) {
    real s = 0.0;
    const long long e = std::min(n, (b + 1) * B);
    for (long long i = b * B; i < e; ++i) s += v[(size_t)i];
    part[(size_t)b] = s;
  }
  real s = 0.0;
  for (real x : part) s += x;
  return s;
}

void StencilOp::resize() {
  const size_t n = (size_t)(three_d ? g.ncell() : g.nsurf());
  for (auto* a : {&diag, &xm, &xp, &ym, &yp, &zm, &zp}) a->assign(n, 0.0);
  fixed.assign(n, 0);
}

void StencilOp::apply(const std::vector<real>& x, std::vector<real>& y) const {
  const int nx = g.nx, ny = g.ny, nz = three_d ? g.nz : 1;
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = three_d ? g.idx(i, j, k) : g.sidx(i, j);
        const long long sy = nx + 2, sz = g.plane();
        real acc = diag[c] * x[c];
        if (i > 0) acc += xm[c] * x[c - 1];
        if (i < nx - 1) acc += xp[c] * x[c + 1];
        if (j > 0) acc += ym[c] * x[c - sy];
        if (j < ny - 1) acc += yp[c] * x[c + sy];
        if (three_d) {
          if (k > 0) acc += zm[c] * x[c - sz];
          if (k < nz - 1) acc += zp[c] * x[c + sz];
        }
        y[c] = acc;
      }
}

namespace {
// Dot product over interior cells with a deterministic blocked reduction.
real dot_interior(const StencilOp& A, const std::vector<real>& a, const std::vector<real>& b,
                  const std::vector<unsigned char>* mask) {
  const int nx = A.g.nx, ny = A.g.ny, nz = A.three_d ? A.g.nz : 1;
  std::vector<real> part((size_t)nz * ny, 0.0);
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      real s = 0.0;
      for (int i = 0; i < nx; ++i) {
        const long long c = A.three_d ? A.g.idx(i, j, k) : A.g.sidx(i, j);
        if (mask && (*mask)[c]) continue;
        s += a[c] * b[c];
      }
      part[(size_t)k * ny + j] = s;
    }
  return det_sum(part);
}
}  // namespace

namespace {
SolveStats pcg_only(const StencilOp& A, const std::vector<real>& b_in, std::vector<real>& x, const SolverCfg& cfg);
SolveStats bicgstab(const StencilOp& A, const std::vector<real>& b_in, std::vector<real>& x, const SolverCfg& cfg);
}  // namespace

// KSP type dispatch (SolverConfig::ksp_type, include/frehg2/linear/SolverConfig.hpp:15): "cg"
// (default) or "bcgs". A CG breakdown is an error, as a diverged KSP is in the reference
// (PetscLinearSolver.cpp:210-218).
SolveStats pcg_solve(const StencilOp& A, const std::vector<real>& b_in, std::vector<real>& x,
                     const SolverCfg& cfg) {
  if (cfg.ksp == 1) return bicgstab(A, b_in, x, cfg);
  return pcg_only(A, b_in, x, cfg);
}

namespace {
SolveStats pcg_only(const StencilOp& A, const std::vector<real>& b_in, std::vector<real>& x,
                    const SolverCfg& cfg) {
  const int nx = A.g.nx, ny = A.g.ny, nz = A.three_d ? A.g.nz : 1;
  const size_t n = b_in.size();
  const long long sy = nx + 2, sz = A.g.plane();
  x.assign(n, 0.0);
  std::vector<real> b = b_in, r(n, 0.0), z(n, 0.0), p(n, 0.0), q(n, 0.0);
  const bool has_fixed = !A.three_d && std::any_of(A.fixed.begin(), A.fixed.end(), [](unsigned char f) { return f != 0; });
  const std::vector<unsigned char>* mask = has_fixed ? &A.fixed : nullptr;

  // ||D^-1 b|| over every row (PETSc's KSP_NORM_PRECONDITIONED at x0 = 0).
  std::vector<real> zb(n, 0.0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = A.three_d ? A.g.idx(i, j, k) : A.g.sidx(i, j);
        zb[c] = b[c] / A.diag[c];
      }
  const real bnorm = std::sqrt(dot_interior(A, zb, zb, nullptr));

  // Decoupled rows (all off-diagonals zero): solved exactly, folded into the neighbours' rhs.
  if (has_fixed) {
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = A.g.sidx(i, j);
        if (A.fixed[c]) x[c] = b[c] / A.diag[c];
      }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = A.g.sidx(i, j);
        if (A.fixed[c]) continue;
        real acc = b[c];
        if (i > 0 && A.fixed[c - 1]) acc -= A.xm[c] * x[c - 1];
        if (i < nx - 1 && A.fixed[c + 1]) acc -= A.xp[c] * x[c + 1];
        if (j > 0 && A.fixed[c - sy]) acc -= A.ym[c] * x[c - sy];
        if (j < ny - 1 && A.fixed[c + sy]) acc -= A.yp[c] * x[c + sy];
        b[c] = acc;
      }
  }
  auto applyFree = [&](const std::vector<real>& v, std::vector<real>& out) {
    if (!has_fixed) {
      A.apply(v, out);
      return;
    }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = A.g.sidx(i, j);
        if (A.fixed[c]) {
          out[c] = 0.0;
          continue;
        }
        real acc = A.diag[c] * v[c];
        if (i > 0 && !A.fixed[c - 1]) acc += A.xm[c] * v[c - 1];
        if (i < nx - 1 && !A.fixed[c + 1]) acc += A.xp[c] * v[c + 1];
        if (j > 0 && !A.fixed[c - sy]) acc += A.ym[c] * v[c - sy];
        if (j < ny - 1 && !A.fixed[c + sy]) acc += A.yp[c] * v[c + sy];
        out[c] = acc;
      }
  };
  (void)sz;

  auto forInterior = [&](auto&& fn) {
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const long long c = A.three_d ? A.g.idx(i, j, k) : A.g.sidx(i, j);
          if (mask && (*mask)[c]) continue;
          fn(c);
        }
  };

  forInterior([&](long long c) {
    r[c] = b[c];
    z[c] = r[c] / A.diag[c];
    p[c] = z[c];
  });
  SolveStats st;
  st.rhs_norm = bnorm;
  const real ttol = std::max(cfg.rtol * bnorm, cfg.atol);
  real rnorm = std::sqrt(dot_interior(A, z, z, mask));
  if (rnorm <= ttol) {
    st.converged = 1;
    st.residual_norm = rnorm;
    return st;
  }
  real rz = dot_interior(A, r, z, mask);
  for (int it = 1; it <= cfg.max_it; ++it) {
    applyFree(p, q);
    const real pq = dot_interior(A, p, q, mask);
    if (!(pq > 0.0) || !std::isfinite(pq)) throw std::runtime_error("oracle pcg: breakdown (p.Ap <= 0)");
    const real alpha = rz / pq;
    forInterior([&](long long c) {
      x[c] += alpha * p[c];
      r[c] -= alpha * q[c];
      z[c] = r[c] / A.diag[c];
    });
    rnorm = std::sqrt(dot_interior(A, z, z, mask));
    st.iterations = it;
    if (rnorm <= ttol) {
      st.converged = 1;
      break;
    }
    const real rz_new = dot_interior(A, r, z, mask);
    const real beta = rz_new / rz;
    rz = rz_new;
    forInterior([&](long long c) { p[c] = z[c] + beta * p[c]; });
  }
  st.residual_norm = rnorm;
  if (!st.converged) throw std::runtime_error("oracle pcg: not converged within max_it");
  return st;
}

// Right-Jacobi-preconditioned BiCGSTAB (van der Vorst 1992), x0 = 0, on the same (possibly
// nonsymmetric) stencil rows, decoupled 2D rows eliminated exactly as in the CG path.
SolveStats bicgstab(const StencilOp& A, const std::vector<real>& b_in, std::vector<real>& x,
                    const SolverCfg& cfg) {
  const int nx = A.g.nx, ny = A.g.ny, nz = A.three_d ? A.g.nz : 1;
  const size_t n = b_in.size();
  const long long sy = nx + 2;
  x.assign(n, 0.0);
  std::vector<real> b = b_in;
  const bool has_fixed = !A.three_d && std::any_of(A.fixed.begin(), A.fixed.end(), [](unsigned char f) { return f != 0; });
  const std::vector<unsigned char>* mask = has_fixed ? &A.fixed : nullptr;
  std::vector<real> zb(n, 0.0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = A.three_d ? A.g.idx(i, j, k) : A.g.sidx(i, j);
        zb[c] = b[c] / A.diag[c];
      }
  const real bnorm = std::sqrt(dot_interior(A, zb, zb, nullptr));
  if (has_fixed) {
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = A.g.sidx(i, j);
        if (A.fixed[c]) x[c] = b[c] / A.diag[c];
      }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = A.g.sidx(i, j);
        if (A.fixed[c]) continue;
        real acc = b[c];
        if (i > 0 && A.fixed[c - 1]) acc -= A.xm[c] * x[c - 1];
        if (i < nx - 1 && A.fixed[c + 1]) acc -= A.xp[c] * x[c + 1];
        if (j > 0 && A.fixed[c - sy]) acc -= A.ym[c] * x[c - sy];
        if (j < ny - 1 && A.fixed[c + sy]) acc -= A.yp[c] * x[c + sy];
        b[c] = acc;
      }
  }
  auto applyFree = [&](const std::vector<real>& v, std::vector<real>& out) {
    if (!has_fixed) {
      A.apply(v, out);
      return;
    }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = A.g.sidx(i, j);
        if (A.fixed[c]) {
          out[c] = 0.0;
          continue;
        }
        real acc = A.diag[c] * v[c];
        if (i > 0 && !A.fixed[c - 1]) acc += A.xm[c] * v[c - 1];
        if (i < nx - 1 && !A.fixed[c + 1]) acc += A.xp[c] * v[c + 1];
        if (j > 0 && !A.fixed[c - sy]) acc += A.ym[c] * v[c - sy];
        if (j < ny - 1 && !A.fixed[c + sy]) acc += A.yp[c] * v[c + sy];
        out[c] = acc;
      }
  };
  auto forInterior = [&](auto&& fn) {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const long long c = A.three_d ? A.g.idx(i, j, k) : A.g.sidx(i, j);
          if (mask && (*mask)[c]) continue;
          fn(c);
        }
  };
  std::vector<real> r(n, 0.0), rh(n, 0.0), p(n, 0.0), v(n, 0.0), y(n, 0.0), s(n, 0.0), zz(n, 0.0), t(n, 0.0),
      w(n, 0.0);
  forInterior([&](long long c) {
    r[c] = b[c];
    rh[c] = b[c];
  });
  SolveStats st;
  st.rhs_norm = bnorm;
  const real ttol = std::max(cfg.rtol * bnorm, cfg.atol);
  auto pnorm = [&](const std::vector<real>& q) {
    forInterior([&](long long c) { w[c] = q[c] / A.diag[c]; });
    return std::sqrt(dot_interior(A, w, w, mask));
  };
  real rn = pnorm(r);
  if (rn <= ttol) {
    st.converged = 1;
    st.residual_norm = rn;
    return st;
  }
  real rho = 1.0, alpha = 1.0, omega = 1.0;
  for (int it = 1; it <= cfg.max_it; ++it) {
    const real rho_new = dot_interior(A, rh, r, mask);
    if (rho_new == 0.0 || !std::isfinite(rho_new)) throw std::runtime_error("oracle bicgstab: breakdown (rho = 0)");
    const real beta = (rho_new / rho) * (alpha / omega);
    rho = rho_new;
    forInterior([&](long long c) {
      p[c] = r[c] + beta * (p[c] - omega * v[c]);
      y[c] = p[c] / A.diag[c];
    });
    applyFree(y, v);
    const real rv = dot_interior(A, rh, v, mask);
    if (rv == 0.0 || !std::isfinite(rv)) throw std::runtime_error("oracle bicgstab: breakdown (rh.v = 0)");
    alpha = rho / rv;
    forInterior([&](long long c) { s[c] = r[c] - alpha * v[c]; });
    st.iterations = it;
    if (pnorm(s) <= ttol) {
      forInterior([&](long long c) { x[c] += alpha * y[c]; });
      st.converged = 1;
      st.residual_norm = pnorm(s);
      return st;
    }
    forInterior([&](long long c) { zz[c] = s[c] / A.diag[c]; });
    applyFree(zz, t);
    const real tt = dot_interior(A, t, t, mask);
    omega = tt > 0.0 ? dot_interior(A, t, s, mask) / tt : 0.0;
    forInterior([&](long long c) {
      x[c] += alpha * y[c] + omega * zz[c];
      r[c] = s[c] - omega * t[c];
    });
    rn = pnorm(r);
    if (rn <= ttol) {
      st.converged = 1;
      st.residual_norm = rn;
      return st;
    }
    if (omega == 0.0) throw std::runtime_error("oracle bicgstab: breakdown (omega = 0)");
  }
  throw std::runtime_error("oracle bicgstab: not converged within max_it");
}
}  // namespace

}  // namespace frehg_oracle
