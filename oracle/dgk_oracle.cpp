// CPU oracle for the tetrahedral DG hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library (libdgk) links,
// loads or calls this file. Only tests/, the smoke check of __graft_entry__.py
// and the CPU-baseline leg of bench.py load it (as liboracle.so through ctypes),
// and then only as the checker or as the timed CPU baseline.
//
// What it restates. The reference (Notargets/DGKernel) is Go and cannot be
// built in this image (no Go toolchain, no gonum, no gocca/libocca), so this
// file restates its CPU path function by function, each citing the reference
// file:line it follows:
//   element/library/gonudg/{jacobi_quadrature,jacobip,rsttoabc,simplex3dp,
//     nodes3d,xyztorst,vandermonde3d,vandermonde2D,lift3d,startup3d,normals3d,
//     ticonnect3d,buildmaps3d}.go      -> element + mesh setup
//   utils/face_connector.go             -> FaceConnector pick/place indices
//   runner/builder/builder.go:113-159   -> aligned partition offsets
//   runner/builder/builder.go:355-410   -> MATMUL / MATMUL_ADD semantics
//   tetnudg/tet_nudg_mesh_physical_derivative_test.go:344-367
//                                        -> calcPhysicalDerivative
//   examples/3d_scalar_burgers_equation.go:415-428,697-717 -> CFL dt, hMin
// and adds the operators the reference only sketches (strong-form
// Lax-Friedrichs flux + LIFT, SSP-RK3, max|u| CFL step), written from
// Hesthaven & Warburton, "Nodal Discontinuous Galerkin Methods" (2008) §2.4,
// §6, §10 and Shu & Osher (1988).
//
// Parity pins (tests/test_oracle_*.py): JacobiGL known values
// (jacobi_quadrature_test.go:9-130), P_0 constants and P000 = sqrt(3)/2
// (jacobip_test.go, simplex3dp_test.go), Dr/Ds/Dt exactness on linears
// (vandermonde3d_test.go:115-159), V*Vinv = I (startup3d_test.go:58-73), face
// areas / outward normals / metric identities (normals3d_test.go), 12 boundary
// faces on the 6-tet cube (ticonnect3d_test.go:110-154), the P=2 closed-form
// derivative on cube-partitioned.neu (tet_nudg_mesh_physical_derivative_test.go:
// 747-919), FaceConnector Verify invariants and element-ID exchange
// (face_connector_test.go:179-397), plus agreement with the independent numpy
// restatement in dgkernel_amd/nudg.py to <= 1e-13 for P <= 6. The LF/LIFT/
// SSP-RK3 operators have no reference implementation: they are pinned by
// invariants (constant state, discrete conservation, third-order convergence
// of SSP-RK3, P+1 convergence of the spatial operator on smooth data).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Vec = std::vector<double>;

// Row-major dense matrix: a[i*c + j] is entry (i, j), like gonum's mat.Dense.
struct Mat {
  int r = 0, c = 0;
  Vec a;
  Mat() = default;
  Mat(int rr, int cc) : r(rr), c(cc), a((size_t)rr * cc, 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)i * c + j]; }
  double operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};

Mat matmul(const Mat& A, const Mat& B) {
  Mat C(A.r, B.c);
  for (int i = 0; i < A.r; ++i)
    for (int k = 0; k < A.c; ++k) {
      double aik = A(i, k);
      for (int j = 0; j < B.c; ++j) C(i, j) += aik * B(k, j);
    }
  return C;
}

Mat transpose(const Mat& A) {
  Mat T(A.c, A.r);
  for (int i = 0; i < A.r; ++i)
    for (int j = 0; j < A.c; ++j) T(j, i) = A(i, j);
  return T;
}

// Inverse by LU with partial pivoting (the method gonum's Dense.Inverse uses;
// vandermonde3d.go:204, startup3d.go:108, lift3d.go:93). Returns false when
// singular.
bool inverse(const Mat& A, Mat& out) {
  int n = A.r;
  Mat LU = A;
  std::vector<int> piv(n);
  for (int i = 0; i < n; ++i) piv[i] = i;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(LU(k, k));
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(LU(i, k)) > best) { best = std::fabs(LU(i, k)); p = i; }
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(LU(k, j), LU(p, j));
      std::swap(piv[k], piv[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      LU(i, k) /= LU(k, k);
      double l = LU(i, k);
      for (int j = k + 1; j < n; ++j) LU(i, j) -= l * LU(k, j);
    }
  }
  out = Mat(n, n);
  for (int col = 0; col < n; ++col) {
    Vec y(n);
    for (int i = 0; i < n; ++i) {
      double s = (piv[i] == col) ? 1.0 : 0.0;
      for (int j = 0; j < i; ++j) s -= LU(i, j) * y[j];
      y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int j = i + 1; j < n; ++j) s -= LU(i, j) * out(j, col);
      out(i, col) = s / LU(i, i);
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Jacobi polynomials and Gauss quadrature
// ---------------------------------------------------------------------------

// Orthonormal Jacobi polynomial P_N^{(alpha,beta)} at x.
// Follows element/library/gonudg/jacobip.go:30-97 step for step.
Vec jacobiP(const Vec& x, double alpha, double beta, int N) {
  size_t Nc = x.size();
  double ab = alpha + beta, ab1 = alpha + beta + 1.0;
  double a1 = alpha + 1.0, b1 = beta + 1.0;
  double gamma0 = std::pow(2.0, ab1) / ab1 * std::tgamma(a1) * std::tgamma(b1) /
                  std::tgamma(ab1);
  Vec P(Nc);
  double sg0 = 1.0 / std::sqrt(gamma0);
  if (N == 0) {
    for (size_t i = 0; i < Nc; ++i) P[i] = sg0;
    return P;
  }
  std::vector<Vec> PL(N + 1, Vec(Nc));
  for (size_t j = 0; j < Nc; ++j) PL[0][j] = sg0;
  double gamma1 = a1 * b1 / (ab + 3.0) * gamma0;
  double sg1 = 1.0 / std::sqrt(gamma1);
  for (size_t j = 0; j < Nc; ++j)
    PL[1][j] = ((ab + 2.0) * x[j] / 2.0 + (alpha - beta) / 2.0) * sg1;
  if (N == 1) return PL[1];
  double aold = 2.0 / (2.0 + ab) * std::sqrt(a1 * b1 / (ab + 3.0));
  for (int i = 1; i <= N - 1; ++i) {
    double h1 = 2.0 * i + ab;
    double fi = (double)i;
    double anew = 2.0 / (h1 + 2.0) *
                  std::sqrt((fi + 1.0) * (fi + ab1) * (fi + a1) * (fi + b1) /
                            (h1 + 1.0) / (h1 + 3.0));
    double bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0);
    for (size_t j = 0; j < Nc; ++j)
      PL[i + 1][j] = 1.0 / anew * (-aold * PL[i - 1][j] + (x[j] - bnew) * PL[i][j]);
    aold = anew;
  }
  return PL[N];
}

// d/dx of the orthonormal Jacobi polynomial; jacobip.go:107-121.
Vec gradJacobiP(const Vec& x, double alpha, double beta, int N) {
  if (N == 0) return Vec(x.size(), 0.0);
  Vec p = jacobiP(x, alpha + 1, beta + 1, N - 1);
  double fN = (double)N;
  double fac = std::sqrt(fN * (fN + alpha + beta + 1));
  for (auto& v : p) v *= fac;
  return p;
}

// Eigenvalues (ascending) of the symmetric tridiagonal matrix with diagonal d
// and off-diagonal e, by Sturm-sequence bisection to full precision. This
// replaces gonum's EigenSym at jacobi_quadrature.go:86-91; the eigenvalues are
// the same numbers up to round-off.
Vec tridiagEigenvalues(const Vec& d, const Vec& e) {
  int n = (int)d.size();
  double lo = std::numeric_limits<double>::max(), hi = -lo;
  for (int i = 0; i < n; ++i) {
    double r = 0.0;
    if (i > 0) r += std::fabs(e[i - 1]);
    if (i < n - 1) r += std::fabs(e[i]);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto countBelow = [&](double x) {
    int cnt = 0;
    double q = 1.0;
    for (int i = 0; i < n; ++i) {
      double off = (i > 0) ? e[i - 1] * e[i - 1] : 0.0;
      q = d[i] - x - ((i > 0) ? off / q : 0.0);
      if (q == 0.0) q = -1e-300;
      if (q < 0) ++cnt;
    }
    return cnt;
  };
  Vec ev(n);
  for (int k = 0; k < n; ++k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      double m = 0.5 * (a + b);
      if (m == a || m == b) break;
      if (countBelow(m) > k) b = m; else a = m;
    }
    ev[k] = 0.5 * (a + b);
  }
  return ev;
}

double gamma0(double alpha, double beta) {  // jacobi_quadrature.go:103-108
  double ab1 = alpha + beta + 1., a1 = alpha + 1., b1 = beta + 1.;
  return std::tgamma(a1) * std::tgamma(b1) * std::pow(2, ab1) / ab1 / std::tgamma(ab1);
}

// Gauss-Jacobi nodes and weights; jacobi_quadrature.go:39-101 (Golub-Welsch).
void jacobiGQ(double alpha, double beta, int N, Vec& X, Vec& W) {
  if (N == 0) {
    X = {-(alpha - beta) / (alpha + beta + 2.)};
    W = {2.};
    return;
  }
  Vec h1(N + 1), d0(N + 1), d1(N);
  for (int i = 0; i <= N; ++i) h1[i] = 2.0 * i + alpha + beta;
  double fac = beta * beta - alpha * alpha;
  for (int i = 0; i <= N; ++i) d0[i] = fac / (h1[i] * (h1[i] + 2.));
  if (alpha + beta < 10 * 1.e-16) d0[0] = 0.;
  for (int i = 0; i < N; ++i) {
    double ip1 = i + 1.0, val = h1[i];
    d1[i] = 2.0 / (val + 2.0) *
            std::sqrt(ip1 * (ip1 + alpha + beta) * (ip1 + alpha) * (ip1 + beta) /
                      (val + 1) / (val + 3));
  }
  X = tridiagEigenvalues(d0, d1);
  W.assign(N + 1, 0.0);
  // First component of each normalized eigenvector from the three-term
  // recurrence of the tridiagonal system (weights = v0^2 * Gamma0).
  for (int k = 0; k <= N; ++k) {
    Vec p(N + 1);
    p[0] = 1.0;
    if (N >= 1) p[1] = (X[k] - d0[0]) / d1[0];
    for (int i = 1; i < N; ++i) p[i + 1] = ((X[k] - d0[i]) * p[i] - d1[i - 1] * p[i - 1]) / d1[i];
    double nrm = 0.0;
    for (double v : p) nrm += v * v;
    W[k] = (1.0 / nrm) * gamma0(alpha, beta);
  }
}

// Gauss-Lobatto nodes; jacobi_quadrature.go:16-37.
Vec jacobiGL(double alpha, double beta, int N) {
  if (N == 0) return {0.0};
  if (N == 1) return {-1.0, 1.0};
  Vec xint, w;
  jacobiGQ(alpha + 1, beta + 1, N - 2, xint, w);
  Vec x(N + 1);
  x[0] = -1.0;
  for (int i = 0; i < N - 1; ++i) x[1 + i] = xint[i];
  x[N] = 1.0;
  return x;
}

// ---------------------------------------------------------------------------
// Collapsed coordinates and the orthonormal simplex basis
// ---------------------------------------------------------------------------

// rsttoabc.go:7-34
void rstToAbc(const Vec& r, const Vec& s, const Vec& t, Vec& a, Vec& b, Vec& c) {
  size_t n = r.size();
  a.resize(n); b.resize(n); c.resize(n);
  const double tol = 1e-12;
  for (size_t i = 0; i < n; ++i) {
    a[i] = (std::fabs(s[i] + t[i]) < tol) ? -1.0 : 2.0 * (1.0 + r[i]) / (-s[i] - t[i]) - 1.0;
    b[i] = (std::fabs(t[i] - 1.0) < tol) ? -1.0 : 2.0 * (1.0 + s[i]) / (1.0 - t[i]) - 1.0;
    c[i] = t[i];
  }
}

// simplex3dp.go:15-36
Vec simplex3DP(const Vec& a, const Vec& b, const Vec& c, int i, int j, int k) {
  Vec h1 = jacobiP(a, 0.0, 0.0, i);
  Vec h2 = jacobiP(b, 2.0 * i + 1, 0.0, j);
  Vec h3 = jacobiP(c, 2.0 * (i + j) + 2, 0.0, k);
  double normConst = 2.0 * std::sqrt(2.0);
  Vec P(a.size());
  for (size_t n = 0; n < a.size(); ++n) {
    double tv1 = normConst * h1[n] * h2[n];
    double tv2 = std::pow(1.0 - b[n], (double)i);
    double tv3 = h3[n] * std::pow(1.0 - c[n], (double)(i + j));
    P[n] = tv1 * tv2 * tv3;
  }
  return P;
}

// simplex3dp.go:56-121
void gradSimplex3DP(const Vec& r, const Vec& s, const Vec& t, int id, int jd, int kd,
                    Vec& dr, Vec& ds, Vec& dt) {
  Vec a, b, c;
  rstToAbc(r, s, t, a, b, c);
  Vec fa = jacobiP(a, 0, 0, id), gb = jacobiP(b, 2.0 * id + 1, 0, jd),
      hc = jacobiP(c, 2.0 * (id + jd) + 2, 0, kd);
  Vec dfa = gradJacobiP(a, 0, 0, id), dgb = gradJacobiP(b, 2.0 * id + 1, 0, jd),
      dhc = gradJacobiP(c, 2.0 * (id + jd) + 2, 0, kd);
  double normFactor = std::pow(2.0, (double)(2 * id + jd) + 1.5);
  size_t n = r.size();
  dr.resize(n); ds.resize(n); dt.resize(n);
  for (size_t i = 0; i < n; ++i) {
    double ai = a[i], bi = b[i], ci = c[i];
    double V3Dr = dfa[i] * gb[i] * hc[i];
    if (id > 0) V3Dr *= std::pow(0.5 * (1.0 - bi), (double)(id - 1));
    if (id + jd > 0) V3Dr *= std::pow(0.5 * (1.0 - ci), (double)(id + jd - 1));
    double V3Ds = 0.5 * (1.0 + ai) * V3Dr;
    double tmp = dgb[i] * std::pow(0.5 * (1.0 - bi), (double)id);
    if (id > 0) tmp -= (0.5 * id) * gb[i] * std::pow(0.5 * (1.0 - bi), (double)(id - 1));
    if (id + jd > 0) tmp *= std::pow(0.5 * (1.0 - ci), (double)(id + jd - 1));
    tmp = fa[i] * tmp * hc[i];
    V3Ds += tmp;
    double V3Dt = 0.5 * (1.0 + ai) * V3Dr + 0.5 * (1.0 + bi) * tmp;
    double tmp2 = dhc[i] * std::pow(0.5 * (1.0 - ci), (double)(id + jd));
    if (id + jd > 0)
      tmp2 -= (0.5 * (id + jd)) * hc[i] * std::pow(0.5 * (1.0 - ci), (double)(id + jd - 1));
    tmp2 = fa[i] * gb[i] * tmp2;
    tmp2 *= std::pow(0.5 * (1.0 - bi), (double)id);
    V3Dt += tmp2;
    dr[i] = V3Dr * normFactor;
    ds[i] = V3Ds * normFactor;
    dt[i] = V3Dt * normFactor;
  }
}

// vandermonde3d.go:15-41
Mat vandermonde3D(int N, const Vec& r, const Vec& s, const Vec& t) {
  int Np = (int)r.size();
  int Ncol = (N + 1) * (N + 2) * (N + 3) / 6;
  Mat V(Np, Ncol);
  Vec a, b, c;
  rstToAbc(r, s, t, a, b, c);
  int sk = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j)
      for (int k = 0; k <= N - i - j; ++k) {
        Vec col = simplex3DP(a, b, c, i, j, k);
        for (int n = 0; n < Np; ++n) V(n, sk) = col[n];
        ++sk;
      }
  return V;
}

// vandermonde3d.go:45-72
void gradVandermonde3D(int N, const Vec& r, const Vec& s, const Vec& t, Mat& Vr, Mat& Vs,
                       Mat& Vt) {
  int Np = (int)r.size();
  int Ncol = (N + 1) * (N + 2) * (N + 3) / 6;
  Vr = Mat(Np, Ncol); Vs = Mat(Np, Ncol); Vt = Mat(Np, Ncol);
  int sk = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j)
      for (int k = 0; k <= N - i - j; ++k) {
        Vec dr, ds, dt;
        gradSimplex3DP(r, s, t, i, j, k, dr, ds, dt);
        for (int n = 0; n < Np; ++n) { Vr(n, sk) = dr[n]; Vs(n, sk) = ds[n]; Vt(n, sk) = dt[n]; }
        ++sk;
      }
}

// vandermonde2D.go:9-72 (Simplex2DP with sqrt(2) normalization, RStoAB)
Mat vandermonde2D(int N, const Vec& R, const Vec& S) {
  int Np = (N + 1) * (N + 2) / 2;
  int Nr = (int)R.size();
  Mat V(Nr, Np);
  Vec a(Nr), b(Nr);
  for (int n = 0; n < Nr; ++n) {
    a[n] = (S[n] != 1) ? 2 * (1 + R[n]) / (1 - S[n]) - 1 : -1;
    b[n] = S[n];
  }
  int sk = 0;
  const double sq2 = 1.4142135623730951;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j) {
      Vec h1 = jacobiP(a, 0, 0, i), h2 = jacobiP(b, 2.0 * i + 1, 0, j);
      for (int n = 0; n < Nr; ++n) {
        double tv2 = 1.0;
        if (i > 0) { tv2 = 1 - b[n]; for (int q = 1; q < i; ++q) tv2 *= (1 - b[n]); }
        V(n, sk) = sq2 * h1[n] * h2[n] * tv2;
      }
      ++sk;
    }
  return V;
}

// ---------------------------------------------------------------------------
// Warp & blend nodes (nodes3d.go), transcribed literally. The reference's
// alphastore indexing and evalwarp differ from the textbook for P >= 3; the
// reference's nodes (and so V, Dr, Ds, Dt, LIFT) are pinned only by following
// nodes3d.go exactly, which this does.
// ---------------------------------------------------------------------------

void equiNodes3D(int p, Vec& r, Vec& s, Vec& t) {  // nodes3d.go:177-204
  int Np = (p + 1) * (p + 2) * (p + 3) / 6;
  r.assign(Np, 0); s.assign(Np, 0); t.assign(Np, 0);
  if (p == 0) { r[0] = s[0] = t[0] = -1.0; return; }
  int sk = 0;
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= p - n; ++m)
      for (int l = 0; l <= p - n - m; ++l) {
        r[sk] = -1.0 + 2.0 * l / (double)p;
        s[sk] = -1.0 + 2.0 * m / (double)p;
        t[sk] = -1.0 + 2.0 * n / (double)p;
        ++sk;
      }
}

Vec evalwarp(int p, const Vec& gaussX, const Vec& xnodes) {  // nodes3d.go:290-323
  Vec xeq(p + 1);
  for (int i = 0; i <= p; ++i) xeq[i] = -1.0 + 2.0 * i / (double)p;
  Vec warp(xnodes.size());
  for (size_t i = 0; i < xnodes.size(); ++i) {
    double x = xnodes[i], warpval = 0.0;
    for (int j = 0; j <= p; ++j) {
      double lag = 1.0;
      for (int k = 0; k <= p; ++k)
        if (k != j) lag *= (x - xeq[k]) / (xeq[j] - xeq[k]);
      warpval += lag * (gaussX[j] - xeq[j]);
    }
    warp[i] = warpval * (1.0 - x * x);
  }
  return warp;
}

void evalshift(int p, double pval, const Vec& L1, const Vec& L2, const Vec& L3, Vec& dx,
               Vec& dy) {  // nodes3d.go:221-287
  size_t n = L1.size();
  Vec gaussX = jacobiGL(0, 0, p);
  for (auto& g : gaussX) g = -g;
  Vec tv1(n), tv2(n), tv3(n);
  for (size_t i = 0; i < n; ++i) {
    tv1[i] = L3[i] - L2[i];
    tv2[i] = L1[i] - L3[i];
    tv3[i] = L2[i] - L1[i];
  }
  Vec w1 = evalwarp(p, gaussX, tv1), w2 = evalwarp(p, gaussX, tv2), w3 = evalwarp(p, gaussX, tv3);
  const double TWOPI = 2.0 * M_PI, FOURPI = 4.0 * M_PI;
  dx.resize(n); dy.resize(n);
  for (size_t i = 0; i < n; ++i) {
    double b1 = L2[i] * L3[i], b2 = L1[i] * L3[i], b3 = L1[i] * L2[i];
    double wf1 = w1[i] * 4.0, wf2 = w2[i] * 4.0, wf3 = w3[i] * 4.0;
    double warp1 = b1 * wf1 * (1.0 + pval * pval * L1[i] * L1[i]);
    double warp2 = b2 * wf2 * (1.0 + pval * pval * L2[i] * L2[i]);
    double warp3 = b3 * wf3 * (1.0 + pval * pval * L3[i] * L3[i]);
    dx[i] = 1.0 * warp1 + std::cos(TWOPI / 3.0) * warp2 + std::cos(FOURPI / 3.0) * warp3;
    dy[i] = 0.0 * warp1 + std::sin(TWOPI / 3.0) * warp2 + std::sin(FOURPI / 3.0) * warp3;
  }
}

void nodes3D(int p, Vec& X, Vec& Y, Vec& Z) {  // nodes3d.go:17-174
  static const double alphastore[15] = {0.0000, 0.0000, 0.0000, 0.1002, 1.1332,
                                        1.5608, 1.3413, 1.2577, 1.1603, 1.10153,
                                        0.6080, 0.4523, 0.8856, 0.8717, 0.9655};
  double alpha = (p < 15 && p >= 0) ? alphastore[p] : 1.0;
  int Np = (p + 1) * (p + 2) * (p + 3) / 6;
  const double tol = 1e-10;
  double sqrt3 = std::sqrt(3.0), sqrt6 = std::sqrt(6.0);
  Vec r, s, t;
  equiNodes3D(p, r, s, t);
  Vec L1(Np), L2(Np), L3(Np), L4(Np);
  for (int i = 0; i < Np; ++i) {
    L1[i] = (1.0 + t[i]) / 2.0;
    L2[i] = (1.0 + s[i]) / 2.0;
    L3[i] = -(1.0 + r[i] + s[i] + t[i]) / 2.0;
    L4[i] = (1.0 + r[i]) / 2.0;
  }
  double v1[3] = {-1.0, -1.0 / sqrt3, -1.0 / sqrt6}, v2[3] = {1.0, -1.0 / sqrt3, -1.0 / sqrt6},
         v3[3] = {0.0, 2.0 / sqrt3, -1.0 / sqrt6}, v4[3] = {0.0, 0.0, 3.0 / sqrt6};
  double t1[4][3], t2[4][3];
  for (int d = 0; d < 3; ++d) {
    t1[0][d] = v2[d] - v1[d]; t2[0][d] = v3[d] - 0.5 * (v1[d] + v2[d]);
    t1[1][d] = v2[d] - v1[d]; t2[1][d] = v4[d] - 0.5 * (v1[d] + v2[d]);
    t1[2][d] = v3[d] - v2[d]; t2[2][d] = v4[d] - 0.5 * (v2[d] + v3[d]);
    t1[3][d] = v3[d] - v1[d]; t2[3][d] = v4[d] - 0.5 * (v1[d] + v3[d]);
  }
  for (int n = 0; n < 4; ++n) {
    double n1 = std::sqrt(t1[n][0] * t1[n][0] + t1[n][1] * t1[n][1] + t1[n][2] * t1[n][2]);
    double n2 = std::sqrt(t2[n][0] * t2[n][0] + t2[n][1] * t2[n][1] + t2[n][2] * t2[n][2]);
    for (int d = 0; d < 3; ++d) { t1[n][d] /= n1; t2[n][d] /= n2; }
  }
  X.assign(Np, 0); Y.assign(Np, 0); Z.assign(Np, 0);
  for (int i = 0; i < Np; ++i) {
    X[i] = L3[i] * v1[0] + L4[i] * v2[0] + L2[i] * v3[0] + L1[i] * v4[0];
    Y[i] = L3[i] * v1[1] + L4[i] * v2[1] + L2[i] * v3[1] + L1[i] * v4[1];
    Z[i] = L3[i] * v1[2] + L4[i] * v2[2] + L2[i] * v3[2] + L1[i] * v4[2];
  }
  Vec shX(Np, 0.0), shY(Np, 0.0), shZ(Np, 0.0);
  for (int face = 0; face < 4; ++face) {
    const Vec *La, *Lb, *Lc, *Ld;
    switch (face) {
      case 0: La = &L1; Lb = &L2; Lc = &L3; Ld = &L4; break;
      case 1: La = &L2; Lb = &L1; Lc = &L3; Ld = &L4; break;
      case 2: La = &L3; Lb = &L1; Lc = &L4; Ld = &L2; break;
      default: La = &L4; Lb = &L1; Lc = &L3; Ld = &L2; break;
    }
    Vec warp1, warp2;
    evalshift(p, alpha, *Lb, *Lc, *Ld, warp1, warp2);  // WarpShiftFace3D, nodes3d.go:213-218
    Vec blend(Np), denom(Np);
    for (int i = 0; i < Np; ++i) {
      blend[i] = (*Lb)[i] * (*Lc)[i] * (*Ld)[i];
      denom[i] = ((*Lb)[i] + 0.5 * (*La)[i]) * ((*Lc)[i] + 0.5 * (*La)[i]) *
                 ((*Ld)[i] + 0.5 * (*La)[i]);
    }
    for (int i = 0; i < Np; ++i)
      if (denom[i] > tol)
        blend[i] = (1.0 + alpha * alpha * (*La)[i] * (*La)[i]) * blend[i] / denom[i];
    for (int i = 0; i < Np; ++i) {
      shX[i] += blend[i] * warp1[i] * t1[face][0] + blend[i] * warp2[i] * t2[face][0];
      shY[i] += blend[i] * warp1[i] * t1[face][1] + blend[i] * warp2[i] * t2[face][1];
      shZ[i] += blend[i] * warp1[i] * t1[face][2] + blend[i] * warp2[i] * t2[face][2];
    }
    for (int i = 0; i < Np; ++i) {
      if ((*La)[i] < tol) {
        int count = 0;
        if ((*Lb)[i] > tol) ++count;
        if ((*Lc)[i] > tol) ++count;
        if ((*Ld)[i] > tol) ++count;
        if (count < 3) {
          shX[i] = warp1[i] * t1[face][0] + warp2[i] * t2[face][0];
          shY[i] = warp1[i] * t1[face][1] + warp2[i] * t2[face][1];
          shZ[i] = warp1[i] * t1[face][2] + warp2[i] * t2[face][2];
        }
      }
    }
  }
  for (int i = 0; i < Np; ++i) { X[i] += shX[i]; Y[i] += shY[i]; Z[i] += shZ[i]; }
}

// xyztorst.go:8-69 (equilateral -> reference tetrahedron)
void xyzToRst(const Vec& X, const Vec& Y, const Vec& Z, Vec& r, Vec& s, Vec& t) {
  double sqrt3 = std::sqrt(3.0), sqrt6 = std::sqrt(6.0);
  double v1[3] = {-1.0, -1.0 / sqrt3, -1.0 / sqrt6}, v2[3] = {1.0, -1.0 / sqrt3, -1.0 / sqrt6},
         v3[3] = {0.0, 2.0 / sqrt3, -1.0 / sqrt6}, v4[3] = {0.0, 0.0, 3.0 / sqrt6};
  double A[3][3];
  for (int d = 0; d < 3; ++d) {
    A[d][0] = 0.5 * (v2[d] - v1[d]);
    A[d][1] = 0.5 * (v3[d] - v1[d]);
    A[d][2] = 0.5 * (v4[d] - v1[d]);
  }
  double off[3];
  for (int d = 0; d < 3; ++d) off[d] = 0.5 * (v2[d] + v3[d] + v4[d] - v1[d]);
  double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
               A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
               A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  double inv[3][3];
  inv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
  inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  inv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
  inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  inv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
  inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
  size_t n = X.size();
  r.resize(n); s.resize(n); t.resize(n);
  for (size_t i = 0; i < n; ++i) {
    double rhs[3] = {X[i] - off[0], Y[i] - off[1], Z[i] - off[2]};
    r[i] = inv[0][0] * rhs[0] + inv[0][1] * rhs[1] + inv[0][2] * rhs[2];
    s[i] = inv[1][0] * rhs[0] + inv[1][1] * rhs[1] + inv[1][2] * rhs[2];
    t[i] = inv[2][0] * rhs[0] + inv[2][1] * rhs[1] + inv[2][2] * rhs[2];
  }
}

// ---------------------------------------------------------------------------
// Reference element (startup3d.go:94-118, 179-209; lift3d.go)
// ---------------------------------------------------------------------------
struct Element {
  int N = 0, Np = 0, Nfp = 0;
  Vec r, s, t;
  Mat V, Vinv, Mass, Dr, Ds, Dt, LIFT;
  std::vector<std::vector<int>> Fmask;  // [4][Nfp]
};

bool buildElement(int N, Element& el) {
  el.N = N;
  el.Np = (N + 1) * (N + 2) * (N + 3) / 6;
  el.Nfp = (N + 1) * (N + 2) / 2;
  Vec x1, y1, z1;
  nodes3D(N, x1, y1, z1);
  xyzToRst(x1, y1, z1, el.r, el.s, el.t);
  el.V = vandermonde3D(N, el.r, el.s, el.t);
  if (!inverse(el.V, el.Vinv)) return false;
  el.Mass = matmul(transpose(el.Vinv), el.Vinv);  // startup3d.go:114-115
  Mat Vr, Vs, Vt;                                  // Dmatrices3D, vandermonde3d.go:76-99
  gradVandermonde3D(N, el.r, el.s, el.t, Vr, Vs, Vt);
  el.Dr = matmul(Vr, el.Vinv);
  el.Ds = matmul(Vs, el.Vinv);
  el.Dt = matmul(Vt, el.Vinv);
  const double NODETOL = 1e-7;  // startup3d.go:81
  el.Fmask.assign(4, {});
  for (int i = 0; i < el.Np; ++i) if (std::fabs(1.0 + el.t[i]) < NODETOL) el.Fmask[0].push_back(i);
  for (int i = 0; i < el.Np; ++i) if (std::fabs(1.0 + el.s[i]) < NODETOL) el.Fmask[1].push_back(i);
  for (int i = 0; i < el.Np; ++i)
    if (std::fabs(1.0 + el.r[i] + el.s[i] + el.t[i]) < NODETOL) el.Fmask[2].push_back(i);
  for (int i = 0; i < el.Np; ++i) if (std::fabs(1.0 + el.r[i]) < NODETOL) el.Fmask[3].push_back(i);
  for (int f = 0; f < 4; ++f) if ((int)el.Fmask[f].size() != el.Nfp) return false;
  // Lift3D, lift3d.go:9-90
  Mat Emat(el.Np, 4 * el.Nfp);
  for (int face = 0; face < 4; ++face) {
    Vec fR(el.Nfp), fS(el.Nfp);
    for (int i = 0; i < el.Nfp; ++i) {
      int idx = el.Fmask[face][i];
      switch (face) {
        case 0: fR[i] = el.r[idx]; fS[i] = el.s[idx]; break;
        case 1: fR[i] = el.r[idx]; fS[i] = el.t[idx]; break;
        case 2: fR[i] = el.s[idx]; fS[i] = el.t[idx]; break;
        default: fR[i] = el.s[idx]; fS[i] = el.t[idx]; break;
      }
    }
    Mat VF = vandermonde2D(N, fR, fS);
    Mat VVt = matmul(VF, transpose(VF)), massFace;
    if (!inverse(VVt, massFace)) return false;
    for (int i = 0; i < el.Nfp; ++i)
      for (int j = 0; j < el.Nfp; ++j) Emat(el.Fmask[face][i], face * el.Nfp + j) = massFace(i, j);
  }
  el.LIFT = matmul(el.V, matmul(transpose(el.V), Emat));
  return true;
}

// ---------------------------------------------------------------------------
// Mesh-dependent setup: StartUp3D affine map, GeometricFactors3D, Normals3D,
// ComputeFscale, tiConnect3D, BuildMaps3D. Per-node fields are stored
// element-major, f[k*Np + i], the layout the reference's runner hands to
// kernels (runner/kernel_copy.go:86-110).
// ---------------------------------------------------------------------------
struct MeshData {
  int K = 0;
  Vec X, Y, Z;                                // Np*K
  Vec J, rx, ry, rz, sx, sy, sz, tx, ty, tz;  // Np*K
  Vec nx, ny, nz, sJ, Fscale;                 // 4*Nfp*K
  std::vector<int> EToE, EToF;                // 4*K
  std::vector<int> vmapM, vmapP, mapP;        // 4*Nfp*K
};

// startup3d.go:126-148
void affineMap(const Element& el, int K, const double* VX, const double* VY, const double* VZ,
               const int* EToV, MeshData& m) {
  int Np = el.Np;
  m.K = K;
  m.X.assign((size_t)Np * K, 0); m.Y = m.X; m.Z = m.X;
  for (int k = 0; k < K; ++k) {
    int va = EToV[4 * k], vb = EToV[4 * k + 1], vc = EToV[4 * k + 2], vd = EToV[4 * k + 3];
    for (int i = 0; i < Np; ++i) {
      double r = el.r[i], s = el.s[i], t = el.t[i];
      size_t q = (size_t)k * Np + i;
      m.X[q] = 0.5 * (-(1.0 + r + s + t) * VX[va] + (1.0 + r) * VX[vb] + (1.0 + s) * VX[vc] + (1.0 + t) * VX[vd]);
      m.Y[q] = 0.5 * (-(1.0 + r + s + t) * VY[va] + (1.0 + r) * VY[vb] + (1.0 + s) * VY[vc] + (1.0 + t) * VY[vd]);
      m.Z[q] = 0.5 * (-(1.0 + r + s + t) * VZ[va] + (1.0 + r) * VZ[vb] + (1.0 + s) * VZ[vc] + (1.0 + t) * VZ[vd]);
    }
  }
}

// out[k*rows + i] = sum_j A(i,j) * in[k*cols + j] for every element k: the
// MATMUL_<name> macro of runner/builder/builder.go:367-386 (device-matrix form
// reads A column-major, A[j*rows + i]; the value is the same product).
void applyPerElement(const Mat& A, const double* in, double* out, long K, bool add) {
#pragma omp parallel for schedule(static)
  for (long k = 0; k < K; ++k) {
    const double* u = in + (size_t)k * A.c;
    double* o = out + (size_t)k * A.r;
    for (int i = 0; i < A.r; ++i) {
      double sum = 0.0;
      for (int j = 0; j < A.c; ++j) sum += A(i, j) * u[j];
      if (add) o[i] += sum; else o[i] = sum;
    }
  }
}

// normals3d.go:11-79 (returns false on J <= 0, the reference's error)
bool geometricFactors(const Element& el, MeshData& m) {
  int Np = el.Np;
  long K = m.K;
  size_t n = (size_t)Np * K;
  Vec xr(n), xs(n), xt(n), yr(n), ys(n), yt(n), zr(n), zs(n), zt(n);
  applyPerElement(el.Dr, m.X.data(), xr.data(), K, false);
  applyPerElement(el.Ds, m.X.data(), xs.data(), K, false);
  applyPerElement(el.Dt, m.X.data(), xt.data(), K, false);
  applyPerElement(el.Dr, m.Y.data(), yr.data(), K, false);
  applyPerElement(el.Ds, m.Y.data(), ys.data(), K, false);
  applyPerElement(el.Dt, m.Y.data(), yt.data(), K, false);
  applyPerElement(el.Dr, m.Z.data(), zr.data(), K, false);
  applyPerElement(el.Ds, m.Z.data(), zs.data(), K, false);
  applyPerElement(el.Dt, m.Z.data(), zt.data(), K, false);
  m.J.assign(n, 0); m.rx = m.ry = m.rz = m.sx = m.sy = m.sz = m.tx = m.ty = m.tz = m.J;
  bool ok = true;
  for (size_t q = 0; q < n; ++q) {
    double J = xr[q] * (ys[q] * zt[q] - zs[q] * yt[q]) - yr[q] * (xs[q] * zt[q] - zs[q] * xt[q]) +
               zr[q] * (xs[q] * yt[q] - ys[q] * xt[q]);
    if (J <= 0) ok = false;
    m.J[q] = J;
    m.rx[q] = (ys[q] * zt[q] - zs[q] * yt[q]) / J;
    m.ry[q] = -(xs[q] * zt[q] - zs[q] * xt[q]) / J;
    m.rz[q] = (xs[q] * yt[q] - ys[q] * xt[q]) / J;
    m.sx[q] = -(yr[q] * zt[q] - zr[q] * yt[q]) / J;
    m.sy[q] = (xr[q] * zt[q] - zr[q] * xt[q]) / J;
    m.sz[q] = -(xr[q] * yt[q] - yr[q] * xt[q]) / J;
    m.tx[q] = (yr[q] * zs[q] - zr[q] * ys[q]) / J;
    m.ty[q] = -(xr[q] * zs[q] - zr[q] * xs[q]) / J;
    m.tz[q] = (xr[q] * ys[q] - yr[q] * xs[q]) / J;
  }
  return ok;
}

// normals3d.go:82-207 (Normals3D + ComputeFscale)
bool normalsAndFscale(const Element& el, MeshData& m) {
  int Np = el.Np, Nfp = el.Nfp;
  long K = m.K;
  size_t nf = (size_t)4 * Nfp * K;
  m.nx.assign(nf, 0); m.ny = m.nz = m.sJ = m.Fscale = m.nx;
  for (long k = 0; k < K; ++k)
    for (int f = 0; f < 4; ++f)
      for (int i = 0; i < Nfp; ++i) {
        size_t v = (size_t)k * Np + el.Fmask[f][i];
        size_t q = (size_t)k * 4 * Nfp + f * Nfp + i;
        double x, y, z;
        switch (f) {
          case 0: x = -m.tx[v]; y = -m.ty[v]; z = -m.tz[v]; break;
          case 1: x = -m.sx[v]; y = -m.sy[v]; z = -m.sz[v]; break;
          case 2: x = m.rx[v] + m.sx[v] + m.tx[v]; y = m.ry[v] + m.sy[v] + m.ty[v];
                  z = m.rz[v] + m.sz[v] + m.tz[v]; break;
          default: x = -m.rx[v]; y = -m.ry[v]; z = -m.rz[v]; break;
        }
        double mag = std::sqrt(x * x + y * y + z * z);
        if (mag < 1e-14) return false;
        m.nx[q] = x / mag; m.ny[q] = y / mag; m.nz[q] = z / mag;
        m.sJ[q] = mag * m.J[v];
        m.Fscale[q] = m.sJ[q] / m.J[v];
      }
  return true;
}

// ticonnect3d.go:24-122. EToVc is the vertex numbering used to identify
// shared faces: equal to EToV for an ordinary mesh; for a periodic mesh, the
// caller passes vertex ids with opposite boundary vertices identified.
void tiConnect3D(long K, const int* EToVc, std::vector<int>& EToE, std::vector<int>& EToF) {
  static const int fv[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};
  long Nnodes = 0;
  for (long q = 0; q < 4 * K; ++q) Nnodes = std::max<long>(Nnodes, EToVc[q] + 1);
  struct F { int64_t id; int elem, face; };
  std::vector<F> faces((size_t)4 * K);
  for (long k = 0; k < K; ++k)
    for (int f = 0; f < 4; ++f) {
      int n3[3] = {EToVc[4 * k + fv[f][0]], EToVc[4 * k + fv[f][1]], EToVc[4 * k + fv[f][2]]};
      std::sort(n3, n3 + 3);
      int64_t id = (int64_t)n3[0] * Nnodes * Nnodes + (int64_t)n3[1] * Nnodes + n3[2];
      faces[(size_t)4 * k + f] = {id, (int)k, f};
    }
  std::sort(faces.begin(), faces.end(), [](const F& a, const F& b) { return a.id < b.id; });
  EToE.resize((size_t)4 * K); EToF.resize((size_t)4 * K);
  for (long k = 0; k < K; ++k)
    for (int f = 0; f < 4; ++f) { EToE[4 * k + f] = (int)k; EToF[4 * k + f] = f; }
  for (size_t i = 0; i + 1 < faces.size(); ++i)
    if (faces[i].id == faces[i + 1].id) {
      int e1 = faces[i].elem, f1 = faces[i].face, e2 = faces[i + 1].elem, f2 = faces[i + 1].face;
      EToE[4 * e1 + f1] = e2; EToF[4 * e1 + f1] = f2;
      EToE[4 * e2 + f2] = e1; EToF[4 * e2 + f2] = f1;
    }
}

// buildmaps3d.go:5-153. Node matching is by coordinates with squared
// tolerance NODETOL^2 (buildmaps3d.go:146). Extension for periodic meshes
// (not in the reference): when period[d] > 0, the neighbour face's
// coordinates are first shifted by the whole number of periods that separates
// the two face centroids along axis d.
void buildMaps3D(const Element& el, MeshData& m, const double* period) {
  int Np = el.Np, Nfp = el.Nfp, NF = 4 * Nfp;
  long K = m.K;
  size_t nf = (size_t)NF * K;
  m.vmapM.resize(nf); m.vmapP.resize(nf); m.mapP.resize(nf);
  for (size_t i = 0; i < nf; ++i) m.mapP[i] = (int)i;
  for (long k = 0; k < K; ++k)
    for (int f = 0; f < 4; ++f)
      for (int i = 0; i < Nfp; ++i) m.vmapM[(size_t)k * NF + f * Nfp + i] = el.Fmask[f][i] + (int)k * Np;
  m.vmapP = m.vmapM;
  const double NODETOL = 1e-7;
#pragma omp parallel for schedule(static)
  for (long k1 = 0; k1 < K; ++k1) {
    std::vector<double> D((size_t)Nfp * Nfp);
    for (int f1 = 0; f1 < 4; ++f1) {
      int k2 = m.EToE[4 * k1 + f1], f2 = m.EToF[4 * k1 + f1];
      if (k2 == k1 && f2 == f1) continue;
      if (k2 < 0 || k2 >= K) continue;
      size_t skM = (size_t)k1 * NF, skP = (size_t)k2 * NF;
      double shift[3] = {0, 0, 0};
      const Vec* C[3] = {&m.X, &m.Y, &m.Z};
      for (int d = 0; d < 3; ++d) {
        if (period == nullptr || period[d] <= 0) continue;
        double c1 = 0, c2 = 0;
        for (int i = 0; i < Nfp; ++i) {
          c1 += (*C[d])[m.vmapM[skM + f1 * Nfp + i]];
          c2 += (*C[d])[m.vmapM[skP + f2 * Nfp + i]];
        }
        shift[d] = std::round((c2 - c1) / Nfp / period[d]) * period[d];
      }
      for (int i = 0; i < Nfp; ++i) {
        int vM = m.vmapM[skM + f1 * Nfp + i];
        for (int j = 0; j < Nfp; ++j) {
          int vP = m.vmapM[skP + f2 * Nfp + j];
          double dx = (m.X[vP] - shift[0]) - m.X[vM];
          double dy = (m.Y[vP] - shift[1]) - m.Y[vM];
          double dz = (m.Z[vP] - shift[2]) - m.Z[vM];
          D[(size_t)i * Nfp + j] = dx * dx + dy * dy + dz * dz;
        }
      }
      for (int i = 0; i < Nfp; ++i) {
        double minD = D[(size_t)i * Nfp];
        int minJ = 0;
        for (int j = 1; j < Nfp; ++j)
          if (D[(size_t)i * Nfp + j] < minD) { minD = D[(size_t)i * Nfp + j]; minJ = j; }
        if (minD < NODETOL * NODETOL) {
          m.vmapP[skM + f1 * Nfp + i] = m.vmapM[skP + f2 * Nfp + minJ];
          m.mapP[skM + f1 * Nfp + i] = (int)(skP + f2 * Nfp + minJ);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Strong-form Burgers operator (not in the reference; H&W 2008 §2.4, §6).
// u_t + div(u^2/2 (1,1,1)) = 0.
//   rhs = -div f(u) + LIFT * (Fscale .* (n.f(u-) - n.f*)),
//   n.f* = 1/2 (n.f(u-) + n.f(u+)) - 1/2 alpha |n.(1,1,1)| (u+ - u-),
//   alpha = max(|u-|, |u+|).
// Per-node metrics are used exactly as GeometricFactors3D produces them.
// ---------------------------------------------------------------------------
struct Solver {
  Element el;
  MeshData m;
  Vec Fs, dFr, dFs, dFt, flux;  // scratch
};

void burgersRHS(Solver& S, const double* u, double* rhs) {
  const Element& el = S.el;
  MeshData& m = S.m;
  int Np = el.Np, Nfp = el.Nfp, NF = 4 * Nfp;
  long K = m.K;
  size_t n = (size_t)Np * K;
  S.Fs.resize(n); S.dFr.resize(n); S.dFs.resize(n); S.dFt.resize(n);
  S.flux.resize((size_t)NF * K);
  for (size_t q = 0; q < n; ++q) S.Fs[q] = 0.5 * u[q] * u[q];
  applyPerElement(el.Dr, S.Fs.data(), S.dFr.data(), K, false);
  applyPerElement(el.Ds, S.Fs.data(), S.dFs.data(), K, false);
  applyPerElement(el.Dt, S.Fs.data(), S.dFt.data(), K, false);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < (long)n; ++q) {
    double Fx = m.rx[q] * S.dFr[q] + m.sx[q] * S.dFs[q] + m.tx[q] * S.dFt[q];
    double Fy = m.ry[q] * S.dFr[q] + m.sy[q] * S.dFs[q] + m.ty[q] * S.dFt[q];
    double Fz = m.rz[q] * S.dFr[q] + m.sz[q] * S.dFs[q] + m.tz[q] * S.dFt[q];
    rhs[q] = -(Fx + Fy + Fz);
  }
#pragma omp parallel for schedule(static)
  for (long q = 0; q < (long)NF * K; ++q) {
    double uM = u[m.vmapM[q]], uP = u[m.vmapP[q]];
    double sn = m.nx[q] + m.ny[q] + m.nz[q];
    double fM = 0.5 * uM * uM, fP = 0.5 * uP * uP;
    double alpha = std::max(std::fabs(uM), std::fabs(uP));
    double fstar = 0.5 * (fM + fP) * sn - 0.5 * alpha * std::fabs(sn) * (uP - uM);
    S.flux[q] = m.Fscale[q] * (fM * sn - fstar);
  }
  applyPerElement(el.LIFT, S.flux.data(), rhs, K, true);
}

// The same operator with the reference's OCCA-OpenMP schedule: one loop
// iteration per partition (@outer -> omp parallel for), a serial element loop
// inside it (@inner), so at most min(NPART, threads) threads are busy
// (documentation/DGKernel/0 - Introduction/intro.md:42-45). Volume term,
// LF flux and LIFT are evaluated element by element. poff[p] .. poff[p+1] are
// the elements of partition p (contiguous, dense order).
void burgersRHSPartitions(Solver& S, const double* u, double* rhs, const long* poff, int npart) {
  const Element& el = S.el;
  MeshData& m = S.m;
  const int Np = el.Np, Nfp = el.Nfp, NF = 4 * Nfp;
#pragma omp parallel for schedule(dynamic, 1)
  for (int p = 0; p < npart; ++p) {
    std::vector<double> F(Np), fl(NF);
    for (long k = poff[p]; k < poff[p + 1]; ++k) {
      const size_t b = (size_t)k * Np;
      for (int i = 0; i < Np; ++i) F[i] = 0.5 * u[b + i] * u[b + i];
      for (int i = 0; i < Np; ++i) {
        double dr = 0.0, ds = 0.0, dt = 0.0;
        for (int j = 0; j < Np; ++j) {
          dr += el.Dr(i, j) * F[j];
          ds += el.Ds(i, j) * F[j];
          dt += el.Dt(i, j) * F[j];
        }
        const size_t q = b + i;
        const double Fx = m.rx[q] * dr + m.sx[q] * ds + m.tx[q] * dt;
        const double Fy = m.ry[q] * dr + m.sy[q] * ds + m.ty[q] * dt;
        const double Fz = m.rz[q] * dr + m.sz[q] * ds + m.tz[q] * dt;
        rhs[q] = -(Fx + Fy + Fz);
      }
      for (int f = 0; f < NF; ++f) {
        const size_t q = (size_t)k * NF + f;
        const double uM = u[m.vmapM[q]], uP = u[m.vmapP[q]];
        const double sn = m.nx[q] + m.ny[q] + m.nz[q];
        const double fM = 0.5 * uM * uM, fP = 0.5 * uP * uP;
        const double alpha = std::max(std::fabs(uM), std::fabs(uP));
        const double fstar = 0.5 * (fM + fP) * sn - 0.5 * alpha * std::fabs(sn) * (uP - uM);
        fl[f] = m.Fscale[q] * (fM * sn - fstar);
      }
      for (int i = 0; i < Np; ++i) {
        double sum = 0.0;
        for (int j = 0; j < NF; ++j) sum += el.LIFT(i, j) * fl[j];
        rhs[b + i] += sum;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Shock capturing (not in the reference; the BASELINE learned-viscosity
// configuration). Written from:
//  * Persson & Peraire, "Sub-cell shock capturing for discontinuous Galerkin
//    methods", AIAA 2006-112: with orthonormal modal coefficients uh of an
//    element, S = sum_{deg m = N} uh_m^2 / sum_m uh_m^2, s = log10 S, and
//      eps = eps0 * 0                                  s < s0 - kappa
//                   1/2 (1 + sin(pi (s - s0) / (2 kappa)))  |s - s0| <= kappa
//                   1                                  s > s0 + kappa,
//    eps0 = ceps * h * lambda / N, h = (6 vol)^(1/3) (the hex edge for the
//    Kuhn split), lambda = sqrt(3) max|u| (the largest |f'(u).n|);
//  * the BR1 discretisation of div(eps grad u) (Bassi & Rebay 1997; H&W 2008
//    §7.1-7.2), strong form, central fluxes plus a jump penalty:
//      q     = grad u + LIFT(Fscale n (u+ - u-)/2)
//      L_v   = div(eps q) - LIFT(Fscale (n.(s- - s+)/2 + tau (u- - u+))),
//      s     = eps q,  tau = tau0 max(eps-, eps+).
// The modal order of Vandermonde3D (i, j, k nested, degree i + j + k) gives
// the top-degree modes.
// ---------------------------------------------------------------------------
struct SensorParams {
  double s0, kappa, ceps;
};

double ppFraction(double S, const SensorParams& sp) {
  if (!(S > 0.0)) return 0.0;
  const double sv = std::log10(S);
  if (sv < sp.s0 - sp.kappa) return 0.0;
  if (sv > sp.s0 + sp.kappa) return 1.0;
  return 0.5 * (1.0 + std::sin(M_PI * (sv - sp.s0) / (2.0 * sp.kappa)));
}

std::vector<int> modeDegrees(int N) {
  std::vector<int> d;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j)
      for (int k = 0; k <= N - i - j; ++k) d.push_back(i + j + k);
  return d;
}

// eps[K], frac[K] (eps / eps0), eps0[K]; feats[K*Np] (float, may be NULL):
// the modal coefficients divided by their Euclidean norm (0 for u = 0).
void ppSensor(const Solver& S, const double* u, const SensorParams& sp, double* eps, double* frac,
              double* eps0, float* feats) {
  const Element& el = S.el;
  const MeshData& m = S.m;
  const int Np = el.Np, N = el.N;
  const std::vector<int> deg = modeDegrees(N);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)m.K; ++k) {
    const double* uk = u + (size_t)k * Np;
    double tot = 0.0, top = 0.0, umax = 0.0;
    std::vector<double> uh(Np);
    for (int i = 0; i < Np; ++i) {
      double a = 0.0;
      for (int j = 0; j < Np; ++j) a += el.Vinv(i, j) * uk[j];
      uh[i] = a;
      tot += a * a;
      if (deg[i] == N) top += a * a;
      umax = std::max(umax, std::fabs(uk[i]));
    }
    const double Sk = tot > 0.0 ? top / tot : 0.0;
    const double fr = ppFraction(Sk, sp);
    const double vol = (4.0 / 3.0) * m.J[(size_t)k * Np];
    const double h = std::cbrt(6.0 * vol);
    const double e0 = sp.ceps * h * std::sqrt(3.0) * umax / N;
    eps[k] = e0 * fr;
    if (frac) frac[k] = fr;
    if (eps0) eps0[k] = e0;
    if (feats) {
      const double inv = tot > 0.0 ? 1.0 / std::sqrt(tot) : 0.0;
      for (int i = 0; i < Np; ++i) feats[(size_t)k * Np + i] = (float)(uh[i] * inv);
    }
  }
}

// out = L_v(u; eps) (overwritten).
void avRHS(Solver& S, const double* u, const double* eps, double tau0, double* out) {
  const Element& el = S.el;
  MeshData& m = S.m;
  const int Np = el.Np, Nfp = el.Nfp, NF = 4 * Nfp;
  const long K = m.K;
  const size_t n = (size_t)Np * K, nf = (size_t)NF * K;
  Vec dr(n), ds(n), dt(n), q[3], g[3], sg[3];
  for (int d = 0; d < 3; ++d) { q[d].assign(n, 0.0); g[d].assign(nf, 0.0); sg[d].assign(n, 0.0); }
  applyPerElement(el.Dr, u, dr.data(), K, false);
  applyPerElement(el.Ds, u, ds.data(), K, false);
  applyPerElement(el.Dt, u, dt.data(), K, false);
  const Vec* RX[3] = {&m.rx, &m.ry, &m.rz};
  const Vec* SX[3] = {&m.sx, &m.sy, &m.sz};
  const Vec* TX[3] = {&m.tx, &m.ty, &m.tz};
  const Vec* NX[3] = {&m.nx, &m.ny, &m.nz};
  for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i)
      q[d][i] = (*RX[d])[i] * dr[i] + (*SX[d])[i] * ds[i] + (*TX[d])[i] * dt[i];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nf; ++i)
      g[d][i] = m.Fscale[i] * (*NX[d])[i] * 0.5 * (u[m.vmapP[i]] - u[m.vmapM[i]]);
    applyPerElement(el.LIFT, g[d].data(), q[d].data(), K, true);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) sg[d][i] = eps[i / Np] * q[d][i];
  }
  std::fill(out, out + n, 0.0);
  for (int d = 0; d < 3; ++d) {
    applyPerElement(el.Dr, sg[d].data(), dr.data(), K, false);
    applyPerElement(el.Ds, sg[d].data(), ds.data(), K, false);
    applyPerElement(el.Dt, sg[d].data(), dt.data(), K, false);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i)
      out[i] += (*RX[d])[i] * dr[i] + (*SX[d])[i] * ds[i] + (*TX[d])[i] * dt[i];
  }
  Vec fl(nf);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)nf; ++i) {
    const int vM = m.vmapM[i], vP = m.vmapP[i];
    const double eM = eps[vM / Np], eP = eps[vP / Np];
    double ns = 0.0;
    for (int d = 0; d < 3; ++d) ns += (*NX[d])[i] * (sg[d][vM] - sg[d][vP]);
    const double tau = tau0 * std::max(eM, eP);
    fl[i] = -m.Fscale[i] * (0.5 * ns + tau * (u[vM] - u[vP]));
  }
  applyPerElement(el.LIFT, fl.data(), out, K, true);
}

// computeMinimumEdgeLength, examples/3d_scalar_burgers_equation.go:697-717
double minNodeDistance(const Element& el, const MeshData& m) {
  int Np = el.Np;
  double hmin = std::numeric_limits<double>::infinity();
#pragma omp parallel for reduction(min : hmin) schedule(static)
  for (long k = 0; k < m.K; ++k)
    for (int i = 0; i < Np; ++i)
      for (int j = i + 1; j < Np; ++j) {
        size_t a = (size_t)k * Np + i, b = (size_t)k * Np + j;
        double dx = m.X[b] - m.X[a], dy = m.Y[b] - m.Y[a], dz = m.Z[b] - m.Z[a];
        double len = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (len > 1e-12 && len < hmin) hmin = len;
      }
  return hmin;
}

double maxAbs(const double* u, size_t n) {
  double mx = 0.0;
#pragma omp parallel for reduction(max : mx) schedule(static)
  for (long q = 0; q < (long)n; ++q) mx = std::max(mx, std::fabs(u[q]));
  return mx;
}

}  // namespace

// ===========================================================================
// C ABI for the ctypes-driven tests. All arrays are caller-allocated; sizes
// follow from (N, K). Returns 0 on success, nonzero on failure.
// ===========================================================================
extern "C" {

int orc_jacobi_gl(double alpha, double beta, int N, double* x) {
  Vec v = jacobiGL(alpha, beta, N);
  std::memcpy(x, v.data(), v.size() * sizeof(double));
  return 0;
}

int orc_jacobi_gq(double alpha, double beta, int N, double* x, double* w) {
  Vec X, W;
  jacobiGQ(alpha, beta, N, X, W);
  std::memcpy(x, X.data(), X.size() * sizeof(double));
  std::memcpy(w, W.data(), W.size() * sizeof(double));
  return 0;
}

int orc_jacobi_p(const double* x, int n, double alpha, double beta, int N, double* out) {
  Vec v = jacobiP(Vec(x, x + n), alpha, beta, N);
  std::memcpy(out, v.data(), n * sizeof(double));
  return 0;
}

int orc_grad_jacobi_p(const double* x, int n, double alpha, double beta, int N, double* out) {
  Vec v = gradJacobiP(Vec(x, x + n), alpha, beta, N);
  std::memcpy(out, v.data(), n * sizeof(double));
  return 0;
}

int orc_simplex3dp(const double* r, const double* s, const double* t, int n, int i, int j, int k,
                   double* P, double* dr, double* ds, double* dt) {
  Vec R(r, r + n), S(s, s + n), T(t, t + n), a, b, c;
  rstToAbc(R, S, T, a, b, c);
  Vec p = simplex3DP(a, b, c, i, j, k);
  std::memcpy(P, p.data(), n * sizeof(double));
  Vec gr, gs, gt;
  gradSimplex3DP(R, S, T, i, j, k, gr, gs, gt);
  std::memcpy(dr, gr.data(), n * sizeof(double));
  std::memcpy(ds, gs.data(), n * sizeof(double));
  std::memcpy(dt, gt.data(), n * sizeof(double));
  return 0;
}

int orc_nodes3d(int N, double* X, double* Y, double* Z) {
  Vec x, y, z;
  nodes3D(N, x, y, z);
  size_t n = x.size();
  std::memcpy(X, x.data(), n * 8); std::memcpy(Y, y.data(), n * 8); std::memcpy(Z, z.data(), n * 8);
  return 0;
}

// Reference element: r,s,t [Np]; V, Vinv, Mass, Dr, Ds, Dt [Np*Np] row-major;
// LIFT [Np*4Nfp] row-major; Fmask [4*Nfp].
int orc_element(int N, double* r, double* s, double* t, double* V, double* Vinv, double* Mass,
                double* Dr, double* Ds, double* Dt, double* LIFT, int* Fmask) {
  Element el;
  if (!buildElement(N, el)) return 1;
  int Np = el.Np, Nfp = el.Nfp;
  std::memcpy(r, el.r.data(), Np * 8); std::memcpy(s, el.s.data(), Np * 8);
  std::memcpy(t, el.t.data(), Np * 8);
  std::memcpy(V, el.V.a.data(), (size_t)Np * Np * 8);
  std::memcpy(Vinv, el.Vinv.a.data(), (size_t)Np * Np * 8);
  std::memcpy(Mass, el.Mass.a.data(), (size_t)Np * Np * 8);
  std::memcpy(Dr, el.Dr.a.data(), (size_t)Np * Np * 8);
  std::memcpy(Ds, el.Ds.a.data(), (size_t)Np * Np * 8);
  std::memcpy(Dt, el.Dt.a.data(), (size_t)Np * Np * 8);
  std::memcpy(LIFT, el.LIFT.a.data(), (size_t)Np * 4 * Nfp * 8);
  for (int f = 0; f < 4; ++f)
    for (int i = 0; i < Nfp; ++i) Fmask[f * Nfp + i] = el.Fmask[f][i];
  return 0;
}

// Opaque solver handle: StartUp3D on a mesh. EToVc: connectivity numbering
// (periodic identification) or NULL to use EToV. period: [3] or NULL.
void* orc_setup(int N, int Nv, const double* VX, const double* VY, const double* VZ, long K,
                const int* EToV, const int* EToVc, const double* period, int* status) {
  Solver* S = new Solver();
  *status = 0;
  if (!buildElement(N, S->el)) { *status = 1; delete S; return nullptr; }
  (void)Nv;
  affineMap(S->el, (int)K, VX, VY, VZ, EToV, S->m);
  if (!geometricFactors(S->el, S->m)) *status = 2;  // J <= 0: the reference errors here
  if (!normalsAndFscale(S->el, S->m)) *status = 3;
  tiConnect3D(K, EToVc ? EToVc : EToV, S->m.EToE, S->m.EToF);
  buildMaps3D(S->el, S->m, period);
  return S;
}

void orc_free(void* h) { delete static_cast<Solver*>(h); }

// Field accessors. which: 0 X,1 Y,2 Z,3 J,4 rx,5 ry,6 rz,7 sx,8 sy,9 sz,
// 10 tx,11 ty,12 tz (Np*K); 20 nx,21 ny,22 nz,23 sJ,24 Fscale (4Nfp*K).
int orc_get_field(void* h, int which, double* out) {
  Solver* S = static_cast<Solver*>(h);
  MeshData& m = S->m;
  const Vec* src = nullptr;
  switch (which) {
    case 0: src = &m.X; break; case 1: src = &m.Y; break; case 2: src = &m.Z; break;
    case 3: src = &m.J; break; case 4: src = &m.rx; break; case 5: src = &m.ry; break;
    case 6: src = &m.rz; break; case 7: src = &m.sx; break; case 8: src = &m.sy; break;
    case 9: src = &m.sz; break; case 10: src = &m.tx; break; case 11: src = &m.ty; break;
    case 12: src = &m.tz; break; case 20: src = &m.nx; break; case 21: src = &m.ny; break;
    case 22: src = &m.nz; break; case 23: src = &m.sJ; break; case 24: src = &m.Fscale; break;
    default: return 1;
  }
  std::memcpy(out, src->data(), src->size() * 8);
  return 0;
}

// which: 0 EToE, 1 EToF (4K); 2 vmapM, 3 vmapP, 4 mapP (4Nfp*K)
int orc_get_imap(void* h, int which, int* out) {
  Solver* S = static_cast<Solver*>(h);
  const std::vector<int>* src = nullptr;
  switch (which) {
    case 0: src = &S->m.EToE; break; case 1: src = &S->m.EToF; break;
    case 2: src = &S->m.vmapM; break; case 3: src = &S->m.vmapP; break;
    case 4: src = &S->m.mapP; break;
    default: return 1;
  }
  std::memcpy(out, src->data(), src->size() * sizeof(int));
  return 0;
}

// calcPhysicalDerivative, tetnudg/tet_nudg_mesh_physical_derivative_test.go:
// 344-367: Ur = Dr*U, Us = Ds*U, Ut = Dt*U, then the chain rule per node.
// All per-node arrays element-major [k*Np + n].
int orc_phys_deriv(int Np, long K, const double* Dr, const double* Ds, const double* Dt,
                   const double* U, const double* const* metrics /*rx,ry,rz,sx,sy,sz,tx,ty,tz*/,
                   double* DuDx, double* DuDy, double* DuDz) {
  Mat mDr(Np, Np), mDs(Np, Np), mDt(Np, Np);
  std::memcpy(mDr.a.data(), Dr, (size_t)Np * Np * 8);
  std::memcpy(mDs.a.data(), Ds, (size_t)Np * Np * 8);
  std::memcpy(mDt.a.data(), Dt, (size_t)Np * Np * 8);
  size_t n = (size_t)Np * K;
  Vec Ur(n), Us(n), Ut(n);
  applyPerElement(mDr, U, Ur.data(), K, false);
  applyPerElement(mDs, U, Us.data(), K, false);
  applyPerElement(mDt, U, Ut.data(), K, false);
  const double *Rx = metrics[0], *Ry = metrics[1], *Rz = metrics[2], *Sx = metrics[3],
               *Sy = metrics[4], *Sz = metrics[5], *Tx = metrics[6], *Ty = metrics[7],
               *Tz = metrics[8];
#pragma omp parallel for schedule(static)
  for (long q = 0; q < (long)n; ++q) {
    DuDx[q] = Rx[q] * Ur[q] + Sx[q] * Us[q] + Tx[q] * Ut[q];
    DuDy[q] = Ry[q] * Ur[q] + Sy[q] * Us[q] + Ty[q] * Ut[q];
    DuDz[q] = Rz[q] * Ur[q] + Sz[q] * Us[q] + Tz[q] * Ut[q];
  }
  return 0;
}

// MATMUL_<X>(IN, OUT, K) / MATMUL_ADD_<X>: runner/builder/builder.go:355-410.
// A is rows x cols row-major.
int orc_matmul(int rows, int cols, const double* A, long K, const double* in, double* out,
               int add) {
  Mat M(rows, cols);
  std::memcpy(M.a.data(), A, (size_t)rows * cols * 8);
  applyPerElement(M, in, out, K, add != 0);
  return 0;
}

// CalculateAlignedOffsetsAndSize, runner/builder/builder.go:113-159.
// offsets[NPART+1] in values; returns total bytes.
long orc_aligned_offsets(int npart, const int* Kp, long specSizeBytes, int dataType, int alignment,
                         long* offsets) {
  long total = 0;
  for (int i = 0; i < npart; ++i) total += Kp[i];
  long bytesPerElement = specSizeBytes / total;
  long valueSize = (dataType == 1 || dataType == 3) ? 4 : 8;
  long valuesPerElement = bytesPerElement / valueSize;
  long a = alignment == 0 ? 1 : alignment;
  long cur = 0;
  for (int i = 0; i < npart; ++i) {
    if (cur % a != 0) cur = ((cur + a - 1) / a) * a;
    offsets[i] = cur / valueSize;
    cur += (long)Kp[i] * valuesPerElement * valueSize;
  }
  if (cur % a != 0) cur = ((cur + a - 1) / a) * a;
  offsets[npart] = cur / valueSize;
  return offsets[npart] * valueSize;
}

// FaceConnector, utils/face_connector.go:43-187. Outputs, for every ordered
// partition pair (src, dst) in src-major order, the pick list (local node in
// src) and, for (dst, src), the place list (P-buffer slot in dst), each in the
// reference's append order (p -> localElem -> face -> fp).
//   counts[src*NP + dst]     : length of pick[src][dst] (= place[dst][src])
//   pick  [sum counts]       : concatenated pick[src][dst] lists, src-major
//   place [sum counts]       : concatenated place[dst][src] lists, dst-major
// numPartitions = max(EToP)+1 as in face_connector.go:59-64.
int orc_face_connector_count(long K, const int* EToP) {
  int np = 0;
  for (long k = 0; k < K; ++k) np = std::max(np, EToP[k] + 1);
  return np;
}

int orc_face_connector(long K, int Nfaces, int Nfp, int Np, const int* VmapP, const int* EToP,
                       long* counts, int* pick, int* place) {
  if (K <= 0 || Nfaces <= 0 || Nfp <= 0 || Np <= 0) return 1;
  int NP = orc_face_connector_count(K, EToP);
  std::vector<long> epp(NP, 0);
  for (long k = 0; k < K; ++k) epp[EToP[k]]++;
  std::vector<std::vector<long>> l2g(NP);
  std::vector<long> g2l(K);
  for (long g = 0; g < K; ++g) {
    int p = EToP[g];
    g2l[g] = (long)l2g[p].size();
    l2g[p].push_back(g);
  }
  std::vector<std::vector<std::vector<int>>> pk(NP, std::vector<std::vector<int>>(NP)),
      pl(NP, std::vector<std::vector<int>>(NP));
  for (int p = 0; p < NP; ++p)
    for (long le = 0; le < epp[p]; ++le) {
      long ge = l2g[p][le];
      for (int f = 0; f < Nfaces; ++f)
        for (int fp = 0; fp < Nfp; ++fp) {
          long gfp = ge * Nfaces * Nfp + f * Nfp + fp;
          long src = VmapP[gfp];
          long se = src / Np;
          int sp = EToP[se];
          long lse = g2l[se];
          long lsn = lse * Np + src % Np;
          long pos = le * Nfaces * Nfp + f * Nfp + fp;
          pk[sp][p].push_back((int)lsn);
          pl[p][sp].push_back((int)pos);
        }
    }
  long o = 0;
  for (int s = 0; s < NP; ++s)
    for (int d = 0; d < NP; ++d) {
      counts[s * NP + d] = (long)pk[s][d].size();
      for (int v : pk[s][d]) pick[o++] = v;
    }
  o = 0;
  for (int d = 0; d < NP; ++d)
    for (int s = 0; s < NP; ++s)
      for (int v : pl[d][s]) place[o++] = v;
  return 0;
}

// Burgers operator and time stepping on an orc_setup handle.
int orc_burgers_rhs(void* h, const double* u, double* rhs) {
  burgersRHS(*static_cast<Solver*>(h), u, rhs);
  return 0;
}

double orc_min_node_distance(void* h) {
  Solver* S = static_cast<Solver*>(h);
  return minNodeDistance(S->el, S->m);
}

// SSP-RK3 (Shu-Osher) with dt = cfl * hmin / max|u| recomputed every step
// (the reference's dt formula, 3d_scalar_burgers_equation.go:415-428, with
// the peak speed measured instead of assumed). If fixed_dt > 0, that dt is
// used instead. dts[nsteps] receives the dt of each step (may be NULL).
int orc_ssprk3(void* h, double* u, int nsteps, double cfl, double hmin, double fixed_dt,
               double* dts) {
  Solver* S = static_cast<Solver*>(h);
  size_t n = (size_t)S->el.Np * S->m.K;
  Vec u1(n), u2(n), L(n);
  for (int step = 0; step < nsteps; ++step) {
    double dt = fixed_dt > 0 ? fixed_dt : cfl * hmin / maxAbs(u, n);
    if (dts) dts[step] = dt;
    burgersRHS(*S, u, L.data());
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)n; ++q) u1[q] = u[q] + dt * L[q];
    burgersRHS(*S, u1.data(), L.data());
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)n; ++q) u2[q] = 0.75 * u[q] + 0.25 * (u1[q] + dt * L[q]);
    burgersRHS(*S, u2.data(), L.data());
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)n; ++q)
      u[q] = (1.0 / 3.0) * u[q] + (2.0 / 3.0) * (u2[q] + dt * L[q]);
  }
  return 0;
}

int orc_burgers_rhs_partitions(void* h, const double* u, double* rhs, const long* poff,
                               int npart) {
  burgersRHSPartitions(*static_cast<Solver*>(h), u, rhs, poff, npart);
  return 0;
}

// SSP-RK3 as orc_ssprk3 with the partition-parallel schedule of
// burgersRHSPartitions for the operator and the stage updates.
int orc_ssprk3_partitions(void* h, double* u, int nsteps, double cfl, double hmin,
                          const long* poff, int npart, double* dts) {
  Solver* S = static_cast<Solver*>(h);
  const int Np = S->el.Np;
  size_t n = (size_t)Np * S->m.K;
  Vec u1(n), u2(n), L(n);
  auto update = [&](auto&& body) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < npart; ++p)
      for (size_t q = (size_t)poff[p] * Np; q < (size_t)poff[p + 1] * Np; ++q) body(q);
  };
  for (int step = 0; step < nsteps; ++step) {
    const double dt = cfl * hmin / maxAbs(u, n);
    if (dts) dts[step] = dt;
    burgersRHSPartitions(*S, u, L.data(), poff, npart);
    update([&](size_t q) { u1[q] = u[q] + dt * L[q]; });
    burgersRHSPartitions(*S, u1.data(), L.data(), poff, npart);
    update([&](size_t q) { u2[q] = 0.75 * u[q] + 0.25 * (u1[q] + dt * L[q]); });
    burgersRHSPartitions(*S, u2.data(), L.data(), poff, npart);
    update([&](size_t q) { u[q] = (1.0 / 3.0) * u[q] + (2.0 / 3.0) * (u2[q] + dt * L[q]); });
  }
  return 0;
}

// Persson-Peraire sensor per element; frac, eps0, feats may be NULL.
int orc_pp_sensor(void* h, const double* u, double s0, double kappa, double ceps, double* eps,
                  double* frac, double* eps0, float* feats) {
  SensorParams sp{s0, kappa, ceps};
  ppSensor(*static_cast<Solver*>(h), u, sp, eps, frac, eps0, feats);
  return 0;
}

// Viscous operator L_v(u; eps) with penalty tau0 max(eps-, eps+).
int orc_av_rhs(void* h, const double* u, const double* eps, double tau0, double* out) {
  avRHS(*static_cast<Solver*>(h), u, eps, tau0, out);
  return 0;
}

// SSP-RK3 of u_t + div(u^2/2 (1,1,1)) = div(eps grad u) with eps from the
// Persson-Peraire sensor of u at the start of each step (frozen over the
// step; mode 0 runs without viscosity). Per step
//   dt = min(cfl hmin / max|u|, cfl_v hmin^2 / max eps)  (second term only if
//   max eps > 0), clipped so that t does not pass t_end (t_end > 0), and
// steps stop when t reaches t_end or after max_steps. Returns the number of
// steps taken; dts[max_steps] (may be NULL) receives each dt, *t_out the
// final time.
int orc_ssprk3_av(void* h, double* u, int max_steps, double t_end, double cfl, double cfl_v,
                  double hmin, int mode, double s0, double kappa, double ceps, double tau0,
                  double* dts, double* t_out) {
  Solver* S = static_cast<Solver*>(h);
  const size_t n = (size_t)S->el.Np * S->m.K;
  Vec u1(n), u2(n), L(n), Lv(n), eps(S->m.K, 0.0);
  SensorParams sp{s0, kappa, ceps};
  double t = 0.0;
  int step = 0;
  auto rhs = [&](const double* x, double* out) {
    burgersRHS(*S, x, out);
    if (mode != 0) {
      avRHS(*S, x, eps.data(), tau0, Lv.data());
      for (size_t q = 0; q < n; ++q) out[q] += Lv[q];
    }
  };
  for (; step < max_steps; ++step) {
    if (t_end > 0.0 && t >= t_end) break;
    double emax = 0.0;
    if (mode != 0) {
      ppSensor(*S, u, sp, eps.data(), nullptr, nullptr, nullptr);
      for (double e : eps) emax = std::max(emax, e);
    }
    double dt = cfl * hmin / maxAbs(u, n);
    if (emax > 0.0) dt = std::min(dt, cfl_v * hmin * hmin / emax);
    if (t_end > 0.0) dt = std::min(dt, t_end - t);
    if (dts) dts[step] = dt;
    rhs(u, L.data());
    for (size_t q = 0; q < n; ++q) u1[q] = u[q] + dt * L[q];
    rhs(u1.data(), L.data());
    for (size_t q = 0; q < n; ++q) u2[q] = 0.75 * u[q] + 0.25 * (u1[q] + dt * L[q]);
    rhs(u2.data(), L.data());
    for (size_t q = 0; q < n; ++q) u[q] = (1.0 / 3.0) * u[q] + (2.0 / 3.0) * (u2[q] + dt * L[q]);
    t += dt;
  }
  if (t_out) *t_out = t;
  return step;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

}  // extern "C"
