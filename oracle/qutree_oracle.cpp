// qutree_oracle.cpp -- FP64 CPU restatement of the QuTree tree-tensor
// contraction path.  TEST INFRASTRUCTURE ONLY: this file is the checker that
// the HIP engine in qutree_amd/ is compared against, and the CPU baseline
// that bench.py times.  Nothing in the product links or calls it.
//
// Every routine cites the reference file:line whose algorithm it restates
// (paths relative to the QuTree source tree).  Loop orders follow the
// reference where they fix the summation order (LA_lib.f matvec/rhomat).
//
// Pinning (what checks each part independently of this restatement):
//   matvec / matrixTensor1 / rhomat / contraction1: closed forms of
//     tests/test_Tensor.cpp:184-205 and brute-force numpy einsum
//     (tests/test_oracle_node.py); BLAS==loop identity (test_TensorBLAS.cpp:60-137).
//   SparseTree sizes: tests/test_SparseMatrixTree.cpp:41-52,108-116.
//   represent / hole contraction: closed forms of test_SparseMatrixTree.cpp:54-87
//     and dense partial traces computed by numpy (tests/test_oracle_tree.py).
//   rho tree / inverse: test_MatrixTree.cpp:69-128 and dense reduced densities.
//   Apply / LayerDerivative / Expectation: dense H|Psi> (numpy kron) projected
//     into the tangent space, energy = <Psi|H|Psi>.
//   HO-DVR basis: numpy eigh of the position matrix (tests/test_oracle_tree.py).
//   BS integrator / CMF integrator: norm and energy conservation and the
//     exact quadratic-Hamiltonian dynamics bound; the exact reference step
//     sequence is parity unpinned (the reference cannot be built here).
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using cd = std::complex<double>;
using std::size_t;
using std::vector;

// ---------------------------------------------------------------------------
// Shapes, tensors, matrices (TensorShape.cpp:3-40, Matrix_Implementation.h:84-96)
// ---------------------------------------------------------------------------
struct Shape {
  vector<size_t> d;
  vector<size_t> bef, aft;
  size_t total = 0;
  Shape() = default;
  explicit Shape(vector<size_t> dims) : d(std::move(dims)) { init(); }
  void init() {
    bef.assign(d.size(), 1);
    aft.assign(d.size(), 1);
    for (size_t k = 0; k < d.size(); ++k)
      for (size_t i = 0; i < k; ++i) bef[k] *= d[i];
    for (size_t k = 0; k < d.size(); ++k)
      for (size_t i = k + 1; i < d.size(); ++i) aft[k] *= d[i];
    total = d.empty() ? 0 : aft[0] * d[0];
  }
  size_t order() const { return d.size(); }
  size_t lastIdx() const { return d.size() - 1; }
  size_t lastDim() const { return d.back(); }
  size_t lastBefore() const { return bef.back(); }
};

struct Tensor {
  Shape s;
  vector<cd> c;
  Tensor() = default;
  explicit Tensor(const Shape& sh) : s(sh), c(sh.total, cd(0., 0.)) {}
  cd& operator[](size_t i) { return c[i]; }
  const cd& operator[](size_t i) const { return c[i]; }
  cd& operator()(size_t i, size_t n) { return c[n * s.lastBefore() + i]; }
  const cd& operator()(size_t i, size_t n) const { return c[n * s.lastBefore() + i]; }
  void zero() { std::fill(c.begin(), c.end(), cd(0., 0.)); }
};

struct Mat {
  size_t d1 = 0, d2 = 0;
  vector<cd> c;
  Mat() = default;
  Mat(size_t a, size_t b) : d1(a), d2(b), c(a * b, cd(0., 0.)) {}
  cd& operator()(size_t i, size_t j) { return c[j * d1 + i]; }
  const cd& operator()(size_t i, size_t j) const { return c[j * d1 + i]; }
};

static Mat identity(size_t n) {
  Mat I(n, n);
  for (size_t i = 0; i < n; ++i) I(i, i) = 1.;
  return I;
}

static Mat mul(const Mat& A, const Mat& B) {  // Matrix_Implementation.h:463-498 (C = A B)
  Mat C(A.d1, B.d2);
  for (size_t j = 0; j < B.d2; ++j)
    for (size_t k = 0; k < A.d2; ++k)
      for (size_t i = 0; i < A.d1; ++i) C(i, j) += A(i, k) * B(k, j);
  return C;
}

static Mat adjoint(const Mat& A) {
  Mat B(A.d2, A.d1);
  for (size_t j = 0; j < A.d2; ++j)
    for (size_t i = 0; i < A.d1; ++i) B(j, i) = std::conj(A(i, j));
  return B;
}

// ---------------------------------------------------------------------------
// Node kernels
// ---------------------------------------------------------------------------
// LA_lib.f:34-81  matvec: mulpsi(i,k,j) (+)= matrix(k,l) * psi(i,l,j); psi(b,a,c)
void matvec(cd* C, const cd* B, const cd* M, size_t a, size_t b, size_t c, bool add) {
  if (!add)
    for (size_t i = 0; i < a * b * c; ++i) C[i] = 0.;
  if (b != 1) {
    for (size_t j = 0; j < c; ++j)
      for (size_t k = 0; k < a; ++k)
        for (size_t l = 0; l < a; ++l) {
          const cd m = M[l * a + k];
          for (size_t i = 0; i < b; ++i) C[i + k * b + j * a * b] += m * B[i + l * b + j * a * b];
        }
  } else {
    for (size_t j = 0; j < c; ++j)
      for (size_t k = 0; k < a; ++k)
        for (size_t l = 0; l < a; ++l) C[k + j * a] += M[l * a + k] * B[l + j * a];
  }
}

// LA_lib.f:244-291 ctmatvec: mulpsi(i,k,j) (+)= conjg(matrix(l,k)) * psi(i,l,j)
void ctmatvec(cd* C, const cd* B, const cd* M, size_t a, size_t b, size_t c, bool add) {
  if (!add)
    for (size_t i = 0; i < a * b * c; ++i) C[i] = 0.;
  for (size_t j = 0; j < c; ++j)
    for (size_t k = 0; k < a; ++k)
      for (size_t l = 0; l < a; ++l) {
        const cd m = std::conj(M[k * a + l]);
        for (size_t i = 0; i < b; ++i) C[i + k * b + j * a * b] += m * B[i + l * b + j * a * b];
      }
}

// LA_lib.f:104-151 tmatvec: mulpsi(i,k,j) (+)= matrix(l,k) * psi(i,l,j)
void tmatvec(cd* C, const cd* B, const cd* M, size_t a, size_t b, size_t c, bool add) {
  if (!add)
    for (size_t i = 0; i < a * b * c; ++i) C[i] = 0.;
  for (size_t j = 0; j < c; ++j)
    for (size_t k = 0; k < a; ++k)
      for (size_t l = 0; l < a; ++l) {
        const cd m = M[k * a + l];
        for (size_t i = 0; i < b; ++i) C[i + k * b + j * a * b] += m * B[i + l * b + j * a * b];
      }
}

// LA_lib.f:314-379 rmatvec: real mulpsi(i,k,j) (+)= matrix(k,l) * psi(i,l,j)
void rmatvec(double* C, const double* B, const double* M, size_t a, size_t b, size_t c, bool add) {
  if (!add)
    for (size_t i = 0; i < a * b * c; ++i) C[i] = 0.;
  for (size_t j = 0; j < c; ++j)
    for (size_t k = 0; k < a; ++k)
      for (size_t l = 0; l < a; ++l) {
        const double m = M[l * a + k];
        for (size_t i = 0; i < b; ++i) C[i + k * b + j * a * b] += m * B[i + l * b + j * a * b];
      }
}

// LA_lib.f:402-467 trmatvec: real mulpsi(i,k,j) (+)= matrix(l,k) * psi(i,l,j)
void trmatvec(double* C, const double* B, const double* M, size_t a, size_t b, size_t c, bool add) {
  if (!add)
    for (size_t i = 0; i < a * b * c; ++i) C[i] = 0.;
  for (size_t j = 0; j < c; ++j)
    for (size_t k = 0; k < a; ++k)
      for (size_t l = 0; l < a; ++l) {
        const double m = M[k * a + l];
        for (size_t i = 0; i < b; ++i) C[i + k * b + j * a * b] += m * B[i + l * b + j * a * b];
      }
}

// LA_lib.f:718-762 rhomat: matrix(j,i) = sum_{n,m} conjg(bra(n,j,m)) ket(n,i,m).
// Always overwrites (LA_lib.f:729).
void rhomat(const cd* bra, const cd* ket, cd* M, size_t a, size_t b, size_t c) {
  for (size_t i = 0; i < a * a; ++i) M[i] = 0.;
  if (b != 1) {
    for (size_t m = 0; m < c; ++m)
      for (size_t i = 0; i < a; ++i)
        for (size_t j = 0; j < a; ++j) {
          cd s = M[i * a + j];
          for (size_t n = 0; n < b; ++n)
            s += std::conj(bra[n + j * b + m * a * b]) * ket[n + i * b + m * a * b];
          M[i * a + j] = s;
        }
  } else {
    if (c > a) {
      for (size_t i = 0; i < a; ++i)
        for (size_t j = 0; j < a; ++j)
          for (size_t m = 0; m < c; ++m) M[i * a + j] += std::conj(bra[j + m * a]) * ket[i + m * a];
    } else {
      for (size_t m = 0; m < c; ++m)
        for (size_t i = 0; i < a; ++i)
          for (size_t j = 0; j < a; ++j) M[i * a + j] += std::conj(bra[j + m * a]) * ket[i + m * a];
    }
  }
}

// TensorBLAS.cpp:104-122 / Tensor_Implementation.h:534-551 (rectangular, any
// matrix type).  C(b,kc,a) (+)= sum_l M(kc,l) B(b,l,a); M is nC x nB.
template <typename MT>
void matrixTensorRect(cd* C, const MT* M, const cd* B, size_t before, size_t nB, size_t nC,
                      size_t after, bool add) {
  if (!add)
    for (size_t i = 0; i < before * nC * after; ++i) C[i] = 0.;
  for (size_t aft = 0; aft < after; ++aft)
    for (size_t l = 0; l < nB; ++l)
      for (size_t k = 0; k < nC; ++k) {
        const auto m = M[l * nC + k];
        for (size_t bef = 0; bef < before; ++bef)
          C[bef + k * before + aft * before * nC] += m * B[bef + l * before + aft * before * nB];
      }
}

// Tensor_Implementation.h:464-479 contraction1 (rectangular):
// S(j,l) (+)= sum_{b,a} conj(A(b,j,a)) B(b,l,a); S is nA x nB.
void contractionRect(cd* S, const cd* A, const cd* B, size_t before, size_t nA, size_t nB,
                     size_t after, bool accumulate) {
  if (!accumulate)
    for (size_t i = 0; i < nA * nB; ++i) S[i] = 0.;
  for (size_t b = 0; b < before; ++b)
    for (size_t j = 0; j < nA; ++j)
      for (size_t l = 0; l < nB; ++l)
        for (size_t a = 0; a < after; ++a)
          S[l * nA + j] += std::conj(A[b + j * before + a * before * nA]) * B[b + l * before + a * before * nB];
}

// matrixTensor(C, M, B, mode, zero) dispatch (Tensor_Implementation.h:554-577,647-661):
// complex square -> matvec.
void matrixTensor(Tensor& C, const Mat& M, const Tensor& B, size_t mode, bool zero) {
  const size_t before = B.s.bef[mode], after = B.s.aft[mode];
  if (M.d1 == M.d2) {
    matvec(C.c.data(), B.c.data(), M.c.data(), M.d1, before, after, !zero);
  } else {
    matrixTensorRect(C.c.data(), M.c.data(), B.c.data(), before, M.d2, M.d1, after, !zero);
  }
}

// Real matrix on complex tensor -> matrixTensor1 loop (Tensor_Implementation.h:575).
void matrixTensorReal(Tensor& C, const vector<double>& M, size_t n, const Tensor& B, size_t mode,
                      bool zero) {
  const size_t before = B.s.bef[mode], after = B.s.aft[mode];
  matrixTensorRect(C.c.data(), M.data(), B.c.data(), before, n, n, after, !zero);
}

Tensor matrixTensor(const Mat& M, const Tensor& B, size_t mode) {
  Shape s = B.s;
  s.d[mode] = M.d1;
  s.init();
  Tensor C(s);
  matrixTensor(C, M, B, mode, true);
  return C;
}

// contraction(S, A, B, k) (Tensor_Implementation.h:482-531): complex square -> rhomat.
Mat contraction(const Tensor& A, const Tensor& B, size_t k) {
  const size_t before = A.s.bef[k], after = A.s.aft[k];
  Mat S(A.s.d[k], B.s.d[k]);
  if (A.s.d[k] == B.s.d[k]) {
    rhomat(A.c.data(), B.c.data(), S.c.data(), A.s.d[k], before, after);
  } else {
    contractionRect(S.c.data(), A.c.data(), B.c.data(), before, A.s.d[k], B.s.d[k], after, false);
  }
  return S;
}

Mat dotProduct(const Tensor& A, const Tensor& B) { return contraction(A, B, A.s.lastIdx()); }

// multStateAB (Tensor_Implementation.h:733-760): C(i,k) (+)= sum_l M(k,l) B(i,l)
void multStateAB(Tensor& C, const Mat& M, const Tensor& B, bool zero) {
  matvec(C.c.data(), B.c.data(), M.c.data(), M.d1, B.s.lastBefore(), 1, !zero);
}
Tensor multStateAB(const Mat& M, const Tensor& B) {
  Tensor C(B.s);
  multStateAB(C, M, B, true);
  return C;
}

// multStateArTB (Tensor_Implementation.h:763-792): C(i,m) += A(n,m) B(i,n)
Tensor multStateArTB(const Mat& A, const Tensor& B) {
  Tensor C(B.s);
  const size_t dimpart = B.s.lastBefore(), nt = B.s.lastDim();
  for (size_t n = 0; n < nt; ++n)
    for (size_t m = 0; m < nt; ++m)
      for (size_t i = 0; i < dimpart; ++i) C[m * dimpart + i] += A(n, m) * B[n * dimpart + i];
  return C;
}

// singleDotProd (Tensor_Implementation.h:416-447), sequential order.
static cd singleDotProd(const Tensor& A, const Tensor& B, size_t n, size_t m) {
  const size_t np = A.s.lastBefore();
  cd r = 0.;
  for (size_t i = 0; i < np; ++i) r += std::conj(A(i, n)) * B(i, m);
  return r;
}

// gramSchmidt (Tensor_Implementation.h:806-853)
void gramSchmidt(Tensor& A) {
  const size_t maxiter = 15;
  const double conver = 1e-12;
  const size_t nt = A.s.lastDim(), dimpart = A.s.lastBefore();
  for (size_t n = 0; n < nt; ++n) {
    size_t iter = 0;
    double accum = 1.;
    while (accum > conver && iter < maxiter) {
      ++iter;
      accum = 0.;
      for (size_t m = 0; m < n; ++m) {
        cd ov = singleDotProd(A, A, m, n);
        accum += std::abs(ov);
        for (size_t i = 0; i < dimpart; ++i) A(i, n) -= ov * A(i, m);
      }
      cd nrm = singleDotProd(A, A, n, n);
      if (std::abs(nrm) != 0.) {
        nrm = std::sqrt(std::real(nrm));
        for (size_t i = 0; i < dimpart; ++i) A(i, n) /= nrm;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Small dense FP64 linear algebra (stands in for Eigen; Matrix_Implementation.h)
// ---------------------------------------------------------------------------
// Hermitian eigendecomposition by cyclic complex Jacobi rotations; eigenvalues
// ascending, phase convention real(U(0,i)) >= 0 (Matrix_Implementation.h:284-308).
void heev(const Mat& Ain, Mat& U, vector<double>& ev) {
  const size_t n = Ain.d1;
  Mat A = Ain;
  U = identity(n);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0., diag = 0.;
    for (size_t j = 0; j < n; ++j)
      for (size_t i = 0; i < n; ++i) {
        if (i != j) off += std::norm(A(i, j));
        else diag += std::norm(A(i, j));
      }
    if (off <= 1e-32 * std::max(diag, 1e-300) || off == 0.) break;
    for (size_t p = 0; p + 1 < n; ++p)
      for (size_t q = p + 1; q < n; ++q) {
        const cd apq = A(p, q);
        const double aa = std::abs(apq);
        if (aa == 0.) continue;
        const double app = std::real(A(p, p)), aqq = std::real(A(q, q));
        const double theta = (aqq - app) / (2. * aa);
        const double t = (theta >= 0 ? 1. : -1.) / (std::abs(theta) + std::sqrt(theta * theta + 1.));
        const double c = 1. / std::sqrt(t * t + 1.), s = t * c;
        const cd ph = apq / aa;  // e^{i phi}
        // rotation R acts on columns p,q: [c, -s*conj? ] -- G = [[c, s*ph],[ -s*conj(ph), c]]
        for (size_t k = 0; k < n; ++k) {  // A <- A G
          const cd akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * std::conj(ph) * akq;
          A(k, q) = s * ph * akp + c * akq;
        }
        for (size_t k = 0; k < n; ++k) {  // A <- G^H A
          const cd apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * ph * aqk;
          A(q, k) = s * std::conj(ph) * apk + c * aqk;
        }
        for (size_t k = 0; k < n; ++k) {  // U <- U G
          const cd ukp = U(k, p), ukq = U(k, q);
          U(k, p) = c * ukp - s * std::conj(ph) * ukq;
          U(k, q) = s * ph * ukp + c * ukq;
        }
      }
  }
  vector<size_t> idx(n);
  for (size_t i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t a, size_t b) { return std::real(A(a, a)) < std::real(A(b, b)); });
  Mat V(n, n);
  ev.assign(n, 0.);
  for (size_t j = 0; j < n; ++j) {
    ev[j] = std::real(A(idx[j], idx[j]));
    // normalise the phase so that U(0,j) is real and non-negative where it is
    // non-zero (gauge that makes the comparison with Eigen's output stable).
    cd ph = 1.;
    const cd u0 = U(0, idx[j]);
    if (std::abs(u0) > 1e-300) ph = std::conj(u0) / std::abs(u0);
    for (size_t i = 0; i < n; ++i) V(i, j) = U(i, idx[j]) * ph;
    if (std::real(V(0, j)) < 0)
      for (size_t i = 0; i < n; ++i) V(i, j) = -V(i, j);
  }
  U = V;
}

// Gauss-Jordan inverse with partial pivoting (stands in for Eigen inverse(),
// Matrix_Implementation.h:311-328).
Mat inverse(const Mat& Ain) {
  const size_t n = Ain.d1;
  Mat A = Ain, I = identity(n);
  for (size_t col = 0; col < n; ++col) {
    size_t piv = col;
    for (size_t r = col + 1; r < n; ++r)
      if (std::abs(A(r, col)) > std::abs(A(piv, col))) piv = r;
    if (piv != col)
      for (size_t k = 0; k < n; ++k) {
        std::swap(A(col, k), A(piv, k));
        std::swap(I(col, k), I(piv, k));
      }
    const cd d = A(col, col);
    for (size_t k = 0; k < n; ++k) {
      A(col, k) /= d;
      I(col, k) /= d;
    }
    for (size_t r = 0; r < n; ++r) {
      if (r == col) continue;
      const cd f = A(r, col);
      if (f == cd(0., 0.)) continue;
      for (size_t k = 0; k < n; ++k) {
        A(r, k) -= f * A(col, k);
        I(r, k) -= f * I(col, k);
      }
    }
  }
  return I;
}

// Thin Q of a Householder QR of the m x n matrix A (m >= n): qr(Matrixcd)
// (Matrix_Implementation.h:795-806) takes householderQ() * Identity(m, n) of
// Eigen's HouseholderQR.  Eigen's makeHouseholder (and LAPACK zlarfg) choose
// beta = -sign(Re alpha) |(alpha, x)| (negative for Re alpha = 0),
// v = [1; x / (alpha - beta)], tau = (beta - alpha) / beta, and the identity
// reflector when x = 0 and alpha is real; Q = H_0 H_1 ... H_{n-1}.
Mat householderQ(const Mat& Ain) {
  const size_t m = Ain.d1, n = Ain.d2;
  Mat A = Ain;
  vector<cd> tau(n, 0.);
  for (size_t j = 0; j < n; ++j) {
    double s = 0.;
    for (size_t r = j + 1; r < m; ++r) s += std::norm(A(r, j));
    const cd alpha = A(j, j);
    if (s == 0. && std::imag(alpha) == 0.) continue;
    const double nr = std::sqrt(std::norm(alpha) + s);
    const double beta = std::real(alpha) >= 0. ? -nr : nr;
    tau[j] = (beta - alpha) / beta;
    const cd sc = 1. / (alpha - beta);
    for (size_t r = j + 1; r < m; ++r) A(r, j) *= sc;
    A(j, j) = beta;
    // A(j:m, k) <- (I - conj(tau) v v^H) A(j:m, k)
    for (size_t k = j + 1; k < n; ++k) {
      cd w = A(j, k);
      for (size_t r = j + 1; r < m; ++r) w += std::conj(A(r, j)) * A(r, k);
      const cd f = std::conj(tau[j]) * w;
      A(j, k) -= f;
      for (size_t r = j + 1; r < m; ++r) A(r, k) -= A(r, j) * f;
    }
  }
  // Q = H_0 ... H_{n-1} [I; 0], applied right to left
  Mat Q(m, n);
  for (size_t i = 0; i < n; ++i) Q(i, i) = 1.;
  for (size_t jj = n; jj-- > 0;) {
    for (size_t k = 0; k < n; ++k) {
      cd w = Q(jj, k);
      for (size_t r = jj + 1; r < m; ++r) w += std::conj(A(r, jj)) * Q(r, k);
      const cd f = tau[jj] * w;
      Q(jj, k) -= f;
      for (size_t r = jj + 1; r < m; ++r) Q(r, k) -= A(r, jj) * f;
    }
  }
  return Q;
}

// Regularised inverse U diag(1/(l + eps e^{-l/eps})) U^H
// (SpectralDecompositionTree.cpp:39-48, Vector_Implementation.h:266-280,
//  Matrix_Implementation.h:635-649)
Mat regularizedInverse(const Mat& rho, double eps) {
  Mat U;
  vector<double> ev;
  heev(rho, U, ev);
  const size_t n = rho.d1;
  Mat R(n, n);
  for (size_t k = 0; k < n; ++k) {
    const double lr = ev[k] + eps * std::exp(-ev[k] / eps);
    const double li = 1. / lr;
    for (size_t j = 0; j < n; ++j)
      for (size_t i = 0; i < n; ++i) R(i, j) += U(i, k) * li * std::conj(U(j, k));
  }
  return R;
}

// projectOut(A, B) = A - B (B^H B)^{-1} B^H A (Tensor_Implementation.h:874-910)
Tensor projectOut(const Tensor& A, const Tensor& B) {
  const Mat ov = dotProduct(B, B);
  const Mat inv = inverse(ov);
  const Mat dp = dotProduct(B, A);
  const Mat prod = mul(inv, dp);
  const Tensor P = multStateArTB(prod, B);
  Tensor R = A;
  for (size_t i = 0; i < R.c.size(); ++i) R[i] -= P[i];
  return R;
}

// ---------------------------------------------------------------------------
// Tree topology (Tree.cpp:35-115, Node.cpp:101-165,225-239, TreeFactory.cpp)
// ---------------------------------------------------------------------------
struct Leaf {
  size_t dim = 0, type = 0, mode = 0;
  double omega = 1., r0 = 0., wfr0 = 0., wfomega = 1.;
  // HO-DVR basis (HO_Basis.cpp:3-34, DVRBasis.h)
  vector<double> x;      // grid
  vector<double> trafo;  // dim x dim, column-major
  vector<double> kin;    // dim x dim, real
  Mat p;                 // complex
};

struct Node {
  int address = -1;
  int parent = -1;
  int childIdx = 0;
  int layer = 0;
  vector<int> children;  // addresses, in child-index order
  bool bottom = false;
  Leaf leaf;
  Shape shape;
  bool isTop() const { return parent < 0; }
};

struct Tree {
  vector<Node> nodes;      // by address (post-order, last child first)
  vector<int> leafNode;    // mode -> bottom-node address
  int top() const { return (int)nodes.size() - 1; }
};

// Recursive builder node used while parsing / constructing.
struct BNode {
  size_t nstates = 0;
  bool isLeaf = false;
  size_t ldim = 0, ltype = 0, lmode = 0;
  vector<std::unique_ptr<BNode>> ch;
};

static std::unique_ptr<BNode> parseNode(std::istream& is) {
  // Node::initialize (Node.cpp:101-165): "nstates -f" then f children; a child
  // with a non-negative second number is a leaf "dim type mode".
  auto b = std::make_unique<BNode>();
  long nst, f;
  is >> nst >> f;
  if (!is || nst <= 0 || f >= 0) throw std::runtime_error("tree text: bad node header");
  b->nstates = (size_t)nst;
  f = -f;
  for (long i = 0; i < f; ++i) {
    auto mark = is.tellg();
    long a, c;
    is >> a >> c;
    is.seekg(mark);
    if (c < 0) {
      b->ch.emplace_back(parseNode(is));
    } else {
      auto L = std::make_unique<BNode>();
      long dim, type, mode;
      is >> dim >> type >> mode;
      L->isLeaf = true;
      L->ldim = (size_t)dim;
      L->ltype = (size_t)type;
      L->lmode = (size_t)mode;
      b->ch.emplace_back(std::move(L));
    }
  }
  return b;
}

static void linearize(const BNode& b, Tree& t, int parent, int childIdx, int layer, int& counter,
                      vector<std::pair<const BNode*, int>>& order) {
  // post-order, visiting the last child first (Node::nextNode, Node.cpp:225-239)
  Node tmp;
  int my = -1;
  vector<int> kids(b.ch.size(), -1);
  bool bottom = (b.ch.size() == 1 && b.ch[0]->isLeaf);
  if (!bottom) {
    for (int k = (int)b.ch.size() - 1; k >= 0; --k) {
      if (b.ch[k]->isLeaf) throw std::runtime_error("mixed leaf/node children unsupported");
      linearize(*b.ch[k], t, -2, k, layer + 1, counter, order);
      kids[k] = counter - 1;
    }
  }
  my = counter++;
  t.nodes.resize(counter);
  Node& n = t.nodes[my];
  n.address = my;
  n.parent = parent;
  n.childIdx = childIdx;
  n.layer = layer;
  n.bottom = bottom;
  n.children = bottom ? vector<int>() : kids;
  for (int k : n.children) t.nodes[k].parent = my;
  vector<size_t> dims;
  if (bottom) {
    n.leaf.dim = b.ch[0]->ldim;
    n.leaf.type = b.ch[0]->ltype;
    n.leaf.mode = b.ch[0]->lmode;
    dims.push_back(n.leaf.dim);
  } else {
    for (int k : n.children) dims.push_back(t.nodes[k].shape.lastDim());
  }
  dims.push_back(b.nstates);
  n.shape = Shape(dims);
  (void)order;
}

// HO_Basis::initialize (HO_Basis.cpp:3-34)
static void initHO(Leaf& L) {
  const size_t d = L.dim;
  Mat X(d, d);
  for (size_t n = 0; n < d; ++n)
    for (size_t m = 0; m < d; ++m) {
      double v = 0.;
      if (m == n + 1) v = std::sqrt(n + 1.);
      if (m + 1 == n) v = std::sqrt(n * 1.);
      v *= std::sqrt(1. / (2. * L.omega));
      X(m, n) = v;
    }
  Mat U;
  vector<double> ev;
  heev(X, U, ev);
  L.x.resize(d);
  L.trafo.assign(d * d, 0.);
  for (size_t i = 0; i < d; ++i) L.x[i] = ev[i] + L.r0;
  for (size_t j = 0; j < d; ++j)
    for (size_t i = 0; i < d; ++i) L.trafo[j * d + i] = std::real(U(i, j));
  // kinetic energy in FBR (HO_Basis.cpp:89-112), then trafo^T K trafo
  vector<double> K(d * d, 0.);
  for (size_t i = 0; i < d; ++i)
    for (size_t j = 0; j < d; ++j) {
      double v = 0.;
      if (i == j + 2) v = std::sqrt((j + 1.) * (j + 2.));
      if (i == j) v = -(2. * j + 1.);
      if (j >= 2 && i == j - 2) v = std::sqrt(j * (j - 1.));
      v *= (-L.omega / 2.) * 0.5;
      K[i * d + j] = v;  // Kin(j, i)
    }
  auto mulR = [&](const vector<double>& A, const vector<double>& B) {
    vector<double> C(d * d, 0.);
    for (size_t j = 0; j < d; ++j)
      for (size_t k = 0; k < d; ++k)
        for (size_t i = 0; i < d; ++i) C[j * d + i] += A[k * d + i] * B[j * d + k];
    return C;
  };
  auto mulRT = [&](const vector<double>& A, const vector<double>& B) {  // A^T B
    vector<double> C(d * d, 0.);
    for (size_t j = 0; j < d; ++j)
      for (size_t i = 0; i < d; ++i)
        for (size_t k = 0; k < d; ++k) C[j * d + i] += A[i * d + k] * B[j * d + k];
    return C;
  };
  L.kin = mulRT(L.trafo, mulR(K, L.trafo));
  // momentum (HO_Basis.cpp:71-87), then trafo^T p trafo
  Mat P(d, d);
  const cd im(0., 1.);
  for (size_t n = 0; n < d; ++n)
    for (size_t m = 0; m < d; ++m) {
      cd v = 0.;
      if (m == n + 1) v = std::sqrt(n + 1.);
      if (m + 1 == n) v = -std::sqrt(n * 1.);
      P(m, n) = v * im * std::sqrt(L.omega / 2.);
    }
  Mat T(d, d);
  for (size_t j = 0; j < d; ++j)
    for (size_t i = 0; i < d; ++i) T(i, j) = L.trafo[j * d + i];
  L.p = mul(adjoint(T), mul(P, T));
}

// HO_Basis::initSPF (HO_Basis.cpp:36-58)
static void initSPF(const Leaf& L, Tensor& phi) {
  const size_t d = L.dim, nst = phi.s.lastDim();
  for (size_t i = 0; i < d; ++i) {
    const double w = L.trafo[i * d + 0] / std::exp(-0.5 * L.omega * std::pow(L.x[i] - L.r0, 2));
    phi(i, 0) = w * std::exp(-0.5 * L.wfomega * std::pow(L.x[i] - L.wfr0, 2));
  }
  for (size_t n = 1; n < nst; ++n)
    for (size_t i = 0; i < d; ++i) phi(i, n) = L.x[i] * phi(i, n - 1);
  gramSchmidt(phi);
}

// Tree(istream) = root Node::initialize + PhysPar per mode (Tree.cpp:137-152)
Tree readTree(const std::string& text) {
  std::istringstream is(text);
  auto root = parseNode(is);
  Tree t;
  int counter = 0;
  vector<std::pair<const BNode*, int>> order;
  linearize(*root, t, -1, 0, 0, counter, order);
  size_t nleaves = 0;
  for (auto& n : t.nodes)
    if (n.bottom) ++nleaves;
  t.leafNode.assign(nleaves, -1);
  for (auto& n : t.nodes)
    if (n.bottom) {
      if (n.leaf.mode >= nleaves) throw std::runtime_error("tree text: leaf mode out of range");
      t.leafNode[n.leaf.mode] = n.address;
    }
  for (size_t m = 0; m < nleaves; ++m) {
    if (t.leafNode[m] < 0) throw std::runtime_error("tree text: missing mode");
    Leaf& L = t.nodes[t.leafNode[m]].leaf;
    double a = 1., b = 0., c = 0., e = 1.;
    if (is >> a >> b >> c >> e) {
      L.omega = a;
      L.r0 = b;
      L.wfr0 = c;
      L.wfomega = e;
    }
    if (L.type != 0) throw std::runtime_error("only HO-DVR leaves (type 0) are on this path");
    initHO(L);
  }
  return t;
}

// TreeFactory::balancedTree (TreeFactory.cpp:143-202), written out as compact
// text.  dim_inc raises the SPF count per layer counted from the bottom.
std::string balancedTreeText(size_t num_leaves, size_t dim_leaves, size_t dim_nodes, size_t dim_inc,
                             double omega, double r0, double wfr0, double wfomega) {
  struct P {
    std::unique_ptr<BNode> b;
  };
  vector<std::unique_ptr<BNode>> nodes;
  const size_t dim_now = std::min(dim_nodes, dim_leaves);
  for (size_t k = 0; k < num_leaves; ++k) {
    auto n = std::make_unique<BNode>();
    n->nstates = dim_now;
    auto L = std::make_unique<BNode>();
    L->isLeaf = true;
    L->ldim = dim_leaves;
    n->ch.emplace_back(std::move(L));
    nodes.emplace_back(std::move(n));
  }
  auto lastDim = [](const BNode& b) { return b.nstates; };
  while (nodes.size() > 1) {  // partition(nodes, 2, dim_nodes) (TreeFactory.cpp:52-105)
    vector<std::unique_ptr<BNode>> groups;
    size_t nloop = nodes.size() / 2;
    for (size_t k = 0; k < nloop; ++k) {
      auto p = std::make_unique<BNode>();
      size_t prod = 1;
      for (size_t l = 0; l < 2; ++l) {
        prod *= lastDim(*nodes[2 * k + l]);
        p->ch.emplace_back(std::move(nodes[2 * k + l]));
      }
      p->nstates = std::min(prod, dim_nodes);
      groups.emplace_back(std::move(p));
    }
    size_t nrest = nodes.size() % 2;
    if (nloop == 0 && nrest > 0) {
      auto p = std::make_unique<BNode>();
      size_t prod = 1;
      for (size_t l = 0; l < nrest; ++l) {
        prod *= lastDim(*nodes[l]);
        p->ch.emplace_back(std::move(nodes[l]));
      }
      p->nstates = std::min(prod, dim_nodes);
      groups.emplace_back(std::move(p));
    } else {
      for (size_t r = 0; r < nrest; ++r) groups.emplace_back(std::move(nodes[nloop * 2 + r]));
    }
    nodes = std::move(groups);
  }
  BNode& root = *nodes.front();
  root.nstates = 1;
  // layer-dependent SPF counts (TreeFactory.cpp:170-190)
  std::function<int(const BNode&)> depth = [&](const BNode& b) -> int {
    if (b.ch.size() == 1 && b.ch[0]->isLeaf) return 0;
    int m = 0;
    for (auto& c : b.ch) m = std::max(m, depth(*c));
    return m + 1;
  };
  // layer() = distance from root; nlayer = max over nodes
  std::function<void(BNode&, int, int&)> maxLayer = [&](BNode& b, int l, int& mx) {
    mx = std::max(mx, l);
    if (!(b.ch.size() == 1 && b.ch[0]->isLeaf))
      for (auto& c : b.ch) maxLayer(*c, l + 1, mx);
  };
  int nlayer = 0;
  maxLayer(root, 0, nlayer);
  std::function<void(BNode&, int, bool)> setDims = [&](BNode& b, int l, bool top) {
    bool bottom = (b.ch.size() == 1 && b.ch[0]->isLeaf);
    if (!bottom)
      for (auto& c : b.ch) setDims(*c, l + 1, false);
    if (!top && !bottom) {
      size_t lb = 1;
      for (auto& c : b.ch) lb *= c->nstates;
      size_t dim = dim_nodes + (size_t)(nlayer - l) * dim_inc;
      if (dim > lb) dim = lb;
      b.nstates = dim;
    }
  };
  setDims(root, 0, true);
  // expandNodes (TreeFactory.cpp:126-141): a non-bottom, non-top node whose
  // SPF count equals the product of its children's is merged into its parent.
  std::function<void(BNode&)> expand = [&](BNode& b) {
    bool bottom = (b.ch.size() == 1 && b.ch[0]->isLeaf);
    if (bottom) return;
    for (auto& c : b.ch) expand(*c);
    vector<std::unique_ptr<BNode>> nc;
    for (auto& c : b.ch) {
      bool cb = (c->ch.size() == 1 && c->ch[0]->isLeaf);
      size_t lb = 1;
      for (auto& g : c->ch) lb *= g->nstates;
      if (!cb && c->nstates == lb) {
        for (auto& g : c->ch) nc.emplace_back(std::move(g));
      } else {
        nc.emplace_back(std::move(c));
      }
    }
    b.ch = std::move(nc);
  };
  expand(root);
  // leaf modes: bottom nodes in post-order (last child first) get f-1 ... 0
  // (Tree::resetLeafModes, Tree.cpp:35-46)
  int mode = (int)num_leaves - 1;
  std::function<void(BNode&)> assign = [&](BNode& b) {
    bool bottom = (b.ch.size() == 1 && b.ch[0]->isLeaf);
    if (bottom) {
      b.ch[0]->lmode = (size_t)mode--;
      return;
    }
    for (int k = (int)b.ch.size() - 1; k >= 0; --k) assign(*b.ch[k]);
  };
  assign(root);
  std::ostringstream os;
  std::function<void(const BNode&)> write = [&](const BNode& b) {
    os << b.nstates << " -" << b.ch.size() << "\n";
    for (auto& c : b.ch) {
      if (c->isLeaf) os << c->ldim << " 0 " << c->lmode << "\n";
      else write(*c);
    }
  };
  write(root);
  for (size_t m = 0; m < num_leaves; ++m) os << omega << " " << r0 << " " << wfr0 << " " << wfomega << "\n";
  return os.str();
}

// ---------------------------------------------------------------------------
// Operators (LeafFunction / LeafMatrix / MultiLeafOperator / SOP)
// ---------------------------------------------------------------------------
enum LeafKind { kX = 0, kX2 = 1, kKin = 2, kP = 3, kId = 4, kMatrix = 5, kPoly = 6 };

struct LeafOp {
  int kind = kX;
  Mat m;  // for kMatrix
};

struct MLO {
  vector<LeafOp> ops;
  vector<size_t> modes;  // targetLeaves_
  vector<size_t> targetLeaves() const {
    vector<size_t> r;
    for (size_t k : modes)
      if (std::find(r.begin(), r.end(), k) == r.end()) r.push_back(k);
    return r;
  }
};

struct SOP {
  vector<MLO> terms;
  vector<cd> coeff;
};

// LeafOperator::apply for DVR leaves (DVRBasis.cpp:8-58, LeafMatrix.cpp:37-41)
static void applyLeafOp(const LeafOp& op, const Leaf& L, Tensor& out, const Tensor& in) {
  const size_t nst = in.s.lastDim(), d = in.s.lastBefore();
  switch (op.kind) {
    case kX:
      for (size_t n = 0; n < nst; ++n)
        for (size_t i = 0; i < d; ++i) out(i, n) = L.x[i] * in(i, n);
      break;
    case kX2:
      for (size_t n = 0; n < nst; ++n)
        for (size_t i = 0; i < d; ++i) out(i, n) = L.x[i] * L.x[i] * in(i, n);
      break;
    case kKin:
      matrixTensorReal(out, L.kin, d, in, 0, true);
      break;
    case kP: {
      matrixTensor(out, L.p, in, 0, true);
      break;
    }
    case kId:
      out = in;
      break;
    case kMatrix:
    case kPoly:  // LeafMatrix assembled on the DVR grid
      matrixTensor(out, op.m, in, 0, true);
      break;
  }
}

// MultiLeafOperator::apply(Phi, leaf) (MultiLeafOperator.cpp:11-37)
static Tensor mloApply(const MLO& M, Tensor Phi, const Leaf& L) {
  Tensor hPhi(Phi.s);
  bool sw = true;
  for (size_t l = 0; l < M.ops.size(); ++l) {
    if (M.modes[l] != L.mode) continue;
    if (sw) applyLeafOp(M.ops[l], L, hPhi, Phi);
    else applyLeafOp(M.ops[l], L, Phi, hPhi);
    sw = !sw;
  }
  return sw ? Phi : hPhi;
}

// ---------------------------------------------------------------------------
// SparseTree (SparseTree.cpp:33-95)
// ---------------------------------------------------------------------------
struct SparseTree {
  vector<int> nodes;           // addresses, ascending (= map order)
  std::map<int, int> co;       // address -> sparse index
  bool active(int addr) const { return co.count(addr) > 0; }
};

SparseTree makeSparseTree(const vector<size_t>& modes, const Tree& t, bool tail) {
  SparseTree s;
  std::map<int, int> co;
  for (size_t k : modes) {
    int a = t.leafNode.at(k);
    while (true) {
      co[a] = 0;
      if (t.nodes[a].isTop()) break;
      a = t.nodes[a].parent;
    }
  }
  for (auto& e : co) s.nodes.push_back(e.first);
  if (!tail) {
    int m;
    for (m = (int)s.nodes.size() - 1; m > 0; --m) {
      const Node& np = t.nodes[s.nodes[m]];
      size_t na = 0;
      for (int c : np.children)
        if (co.count(c)) ++na;
      if (na > 1) break;
    }
    if (m < (int)s.nodes.size()) ++m;
    s.nodes.resize(m);
  }
  for (size_t i = 0; i < s.nodes.size(); ++i) s.co[s.nodes[i]] = (int)i;
  return s;
}

struct SparseMatrixTree {
  SparseTree st;
  vector<Mat> m;  // by sparse index
  bool active(int a) const { return st.active(a); }
  Mat& operator[](int a) { return m[st.co.at(a)]; }
  const Mat& operator[](int a) const { return m[st.co.at(a)]; }
};

static SparseMatrixTree makeSMT(const SparseTree& st, const Tree& t) {
  SparseMatrixTree s;
  s.st = st;
  for (int a : st.nodes) {
    size_t n = t.nodes[a].shape.lastDim();
    s.m.emplace_back(n, n);
  }
  return s;
}

// ---------------------------------------------------------------------------
// Tree functions
// ---------------------------------------------------------------------------
using TensorTree = vector<Tensor>;
using MatrixTree = vector<Mat>;

TensorTree makeTensorTree(const Tree& t) {
  TensorTree P;
  for (auto& n : t.nodes) P.emplace_back(n.shape);
  return P;
}

// TensorTree::fillRandom (TensorTree_Implementation.h:39-76), generate()
// (Tensor_Extension_Implementation.h:252-257): uniform[-1,1] into the real part.
void fillRandom(TensorTree& P, const Tree& t, std::mt19937& gen, bool delta_lowest) {
  for (auto& n : t.nodes) {
    Tensor& Phi = P[n.address];
    if (n.bottom) {
      initSPF(n.leaf, Phi);
    } else {
      std::uniform_real_distribution<double> dist(-1., 1.);
      for (size_t i = 0; i < Phi.c.size(); ++i) Phi[i] = dist(gen);
      if (delta_lowest) {
        for (size_t i = 0; i < Phi.s.lastBefore(); ++i) Phi(i, 0) = 0.;
        Phi(0, 0) = 1.;
      }
      gramSchmidt(Phi);
    }
  }
}

// TreeFunctions::dotProduct (MatrixTreeFunctions_Implementation.h:16-39)
MatrixTree dotProductTree(const TensorTree& Bra, const TensorTree& Ket, const Tree& t) {
  MatrixTree S(t.nodes.size());
  for (auto& n : t.nodes) {
    Tensor K = Ket[n.address];
    if (!n.bottom)
      for (size_t k = 0; k < n.children.size(); ++k) K = matrixTensor(S[n.children[k]], K, k);
    S[n.address] = contraction(Bra[n.address], K, n.shape.lastIdx());
  }
  return S;
}

// TreeFunctions::contraction(Rho, Psi, tree, true) (MatrixTreeFunctions_Implementation.h:59-104)
MatrixTree rhoTree(const TensorTree& P, const Tree& t) {
  MatrixTree R(t.nodes.size());
  for (int a = t.top(); a >= 0; --a) {
    const Node& n = t.nodes[a];
    if (n.isTop()) {
      R[a] = identity(n.shape.lastDim());
    } else {
      const Tensor& Ap = P[n.parent];
      Tensor K = matrixTensor(R[n.parent], Ap, Ap.s.lastIdx());
      R[a] = contraction(Ap, K, n.childIdx);
    }
  }
  return R;
}

// apply() (SparseMatrixTreeFunctions_Implementation.h:299-352)
static Tensor applyNode(const SparseMatrixTree& mat, const SparseMatrixTree* holes, Tensor Phi,
                        const Tree& t, const Node& node, int skip) {
  Tensor cur = Phi;
  size_t nact = 0;
  if (!node.bottom) {
    for (size_t k = 0; k < node.children.size(); ++k) {
      int c = node.children[k];
      if ((int)k == skip || !mat.active(c)) continue;
      cur = matrixTensor(mat[c], cur, k);
      ++nact;
    }
  }
  (void)nact;
  const int parentIdx = (int)node.children.size() + (node.bottom ? 1 : 0);
  if (!node.isTop() && skip != parentIdx && holes != nullptr) {
    if (mat.st.active(node.address) && holes->active(node.address))
      cur = matrixTensor((*holes)[node.address], cur, node.shape.lastIdx());
  }
  return cur;
}

// represent (SparseMatrixTreeFunctions_Implementation.h:18-80)
void represent(SparseMatrixTree& h, const MLO& M, const TensorTree& P, const Tree& t) {
  for (int a : h.st.nodes) {
    const Node& n = t.nodes[a];
    if (n.isTop()) continue;
    Tensor hK;
    if (n.bottom) {
      hK = mloApply(M, P[a], n.leaf);
    } else {
      hK = applyNode(h, nullptr, P[a], t, n, (int)n.children.size());
    }
    h[a] = contraction(P[a], hK, n.shape.lastIdx());
  }
}

// contraction / contractionLayer (SparseMatrixTreeFunctions_Implementation.h:145-210)
void holeContraction(SparseMatrixTree& holes, const SparseMatrixTree& mats, const TensorTree& P,
                     const Tree& t) {
  for (int i = (int)holes.st.nodes.size() - 1; i >= 0; --i) {
    int a = holes.st.nodes[i];
    const Node& n = t.nodes[a];
    if (n.isTop()) {
      holes[a] = identity(n.shape.lastDim());
      continue;
    }
    const Node& p = t.nodes[n.parent];
    Tensor hK = applyNode(mats, &holes, P[n.parent], t, p, n.childIdx);
    holes[a] = contraction(P[n.parent], hK, n.childIdx);
  }
}

struct HamRep {
  MatrixTree rho, rhoInv;
  vector<SparseMatrixTree> hMats, hCon;
};

HamRep makeHamRep(const SOP& H, const Tree& t) {
  HamRep r;
  for (auto& M : H.terms) {
    SparseTree st = makeSparseTree(M.targetLeaves(), t, true);
    r.hMats.push_back(makeSMT(st, t));
    r.hCon.push_back(makeSMT(st, t));
  }
  return r;
}

// HamiltonianRepresentation::build (HamiltonianRepresentation.h:62-87)
void build(HamRep& r, const SOP& H, const TensorTree& P, const Tree& t, double eps) {
  r.rho = rhoTree(P, t);
  r.rhoInv.assign(t.nodes.size(), Mat());
  // The loops below are independent per node / per term.  The default build
  // runs them on one thread like the reference; the all-core CPU baseline
  // build (-fopenmp) spreads them over threads, with the same arithmetic.
#pragma omp parallel for schedule(dynamic)
  for (long a = 0; a < (long)t.nodes.size(); ++a) r.rhoInv[a] = regularizedInverse(r.rho[a], eps);
#pragma omp parallel for schedule(dynamic)
  for (long l = 0; l < (long)H.terms.size(); ++l) represent(r.hMats[l], H.terms[l], P, t);
#pragma omp parallel for schedule(dynamic)
  for (long l = 0; l < (long)H.terms.size(); ++l) holeContraction(r.hCon[l], r.hMats[l], P, t);
}

// Apply (HamiltonianRepresentation.cpp:6-43)
Tensor Apply(const SOP& H, const Tensor& Phi, const HamRep& r, const Tree& t, const Node& node) {
  Tensor dPhi(Phi.s);
  for (size_t l = 0; l < H.terms.size(); ++l) {
    const auto& hm = r.hMats[l];
    if (!hm.active(node.address)) continue;
    Tensor Psi(Phi.s);
    for (size_t i = 0; i < Psi.c.size(); ++i) Psi[i] = H.coeff[l] * Phi[i];
    // TreeFunctions::apply(mats, Phi, M, node) (SparseMatrixTreeFunctions_Implementation.h:394-404)
    if (node.bottom) Psi = mloApply(H.terms[l], Psi, node.leaf);
    else Psi = applyNode(hm, nullptr, Psi, t, node, (int)node.children.size());
    const auto& hh = r.hCon[l];
    if (!node.isTop() && hh.active(node.address)) {
      multStateAB(dPhi, hh[node.address], Psi, false);
    } else {
      for (size_t i = 0; i < dPhi.c.size(); ++i) dPhi[i] += Psi[i];
    }
  }
  return dPhi;
}

Mat Expectation(const SOP& H, const TensorTree& P, const HamRep& r, const Tree& t) {
  const Node& top = t.nodes[t.top()];
  Tensor HPhi = Apply(H, P[top.address], r, t, top);
  return dotProduct(P[top.address], HPhi);
}

// LayerDerivative (HamiltonianRepresentation.cpp:45-72)
Tensor LayerDerivative(const SOP& H, const Tensor& Phi, const HamRep& r, const Tree& t,
                       const Node& node, cd phase) {
  Tensor dPhi = Apply(H, Phi, r, t, node);
  if (!node.isTop()) {
    dPhi = projectOut(dPhi, Phi);
    dPhi = multStateAB(r.rhoInv[node.address], dPhi);
  }
  const cd f = -phase * cd(0., 1.);
  for (auto& x : dPhi.c) x *= f;
  return dPhi;
}

// Gradient of E = <Psi|H|Psi>/<Psi|Psi> with respect to conj(A_v) at every
// node, for a tree whose non-top nodes are orthonormal (the gauge in which
// HamiltonianRepresentation works):
//   dE/dA_v* = (H_eff,v A_v - E rho_v x_last A_v) / <A_top|A_top>,
//   H_eff,v A_v = Apply(v) + M_v x_last A_v,
// where Apply is HamiltonianRepresentation::Apply (the terms active at v)
// and M_v = sum over the terms inactive at v of c_t hole_t(v), their mean
// field, obtained top-down from the parent p:
//   M_v = contraction(A_p, sum_{t active at p, not at v} c_t apply_t(A_p)
//                          + M_p x_last A_p, childIdx(v)).
// The reference drops the inactive terms because (1 - P) removes them from
// the MCTDH derivative; the energy gradient keeps them.  No reference
// counterpart: this pins the device gradient, and finite differences of
// the energy pin this function.
TensorTree EnergyGradient(const SOP& H, const TensorTree& P, const HamRep& r, const Tree& t, double& E) {
  const Node& top = t.nodes[t.top()];
  const cd num = Expectation(H, P, r, t)(0, 0);
  const cd Nrm = dotProduct(P[top.address], P[top.address])(0, 0);
  E = std::real(num / Nrm);
  MatrixTree M(t.nodes.size());
  for (int a = (int)t.nodes.size() - 1; a >= 0; --a) {  // parents before children
    const Node& n = t.nodes[a];
    if (n.isTop()) continue;
    const Node& p = t.nodes[n.parent];
    const Tensor& Ap = P[p.address];
    Tensor Y(Ap.s);
    for (size_t l = 0; l < H.terms.size(); ++l) {
      if (!r.hMats[l].active(p.address) || r.hMats[l].active(a)) continue;
      Tensor Psi(Ap.s);
      for (size_t i = 0; i < Psi.c.size(); ++i) Psi[i] = H.coeff[l] * Ap[i];
      Psi = applyNode(r.hMats[l], &r.hCon[l], Psi, t, p, n.childIdx);
      for (size_t i = 0; i < Y.c.size(); ++i) Y[i] += Psi[i];
    }
    if (!p.isTop()) {
      Tensor K = matrixTensor(M[p.address], Ap, Ap.s.lastIdx());
      for (size_t i = 0; i < Y.c.size(); ++i) Y[i] += K[i];
    }
    M[a] = contraction(Ap, Y, n.childIdx);
  }
  TensorTree G = makeTensorTree(t);
  for (auto& n : t.nodes) {
    const Tensor& A = P[n.address];
    Tensor g = Apply(H, A, r, t, n);
    if (n.isTop()) {
      for (size_t i = 0; i < g.c.size(); ++i) g[i] -= E * A[i];
    } else {
      Mat X = M[n.address];
      const Mat& rho = r.rho[n.address];
      for (size_t i = 0; i < X.c.size(); ++i) X.c[i] -= E * rho.c[i];
      multStateAB(g, X, A, false);
    }
    for (auto& x : g.c) x /= Nrm;
    G[n.address] = g;
  }
  return G;
}

// ---------------------------------------------------------------- TTNO
// Tree tensor network operator (TensorTreeOperator.h:23-58): one tensor per
// node of the operator tree, bottom (dim^2, D_v) with the leaf matrices as
// columns (toMatrix(B, l), TensorTreeOperator.cpp:40-47), upper
// (D_c1, ..., D_ck, D_v), top (D_c1, ..., 1).
static vector<size_t> indexMapping(size_t L, const Shape& s) {
  vector<size_t> ls(s.d.size());
  for (size_t k = 0; k < s.d.size(); ++k) {
    ls[k] = L % s.d[k];
    L /= s.d[k];
  }
  return ls;
}

static Mat leafColumn(const Tensor& B, size_t l, size_t dim) {
  Mat h(dim, dim);
  for (size_t I = 0; I < dim * dim; ++I) h.c[I] = B(I, l);
  return h;
}

// TTOrepresentation::calculate (TTOrepresentation.h:81-103): bottom-up
// rep[v][l] = <Psi_v| (operator of bond index l) |Chi_v>.
vector<vector<Mat>> ttoRepresentation(const TensorTree& Psi, const vector<Tensor>& H, const TensorTree& Chi,
                                      const Tree& t) {
  vector<vector<Mat>> rep(t.nodes.size());
  for (const Node& node : t.nodes) {
    const int a = node.address;
    const Tensor& B = H[a];
    const size_t D = B.s.d.back(), n = node.shape.d.back();
    rep[a].assign(D, Mat(n, n));
    if (node.bottom) {
      for (size_t l = 0; l < D; ++l) {
        Tensor hChi = matrixTensor(leafColumn(B, l, node.leaf.dim), Chi[a], 0);
        Mat hij = dotProduct(Psi[a], hChi);
        for (size_t i = 0; i < hij.c.size(); ++i) rep[a][l].c[i] += hij.c[i];
      }
    } else {
      for (size_t L = 0; L < B.s.total; ++L) {
        vector<size_t> ls = indexMapping(L, B.s);
        Tensor A = Chi[a];
        for (size_t k = 0; k < node.children.size(); ++k) A = matrixTensor(rep[node.children[k]][ls[k]], A, k);
        for (auto& x : A.c) x *= B[L];
        Mat hij = dotProduct(Psi[a], A);
        for (size_t i = 0; i < hij.c.size(); ++i) rep[a][ls.back()].c[i] += hij.c[i];
      }
    }
  }
  return rep;
}

// TTOcontraction::calculate (TTOcontraction.h:58-112): top-down
// con[v][l] = contraction over the legs of the parent but v's of
// <Psi_p| siblings' rep x B_p x con[p] |Chi_p> with v's bond index l.
vector<vector<Mat>> ttoContraction(const TensorTree& Psi, const vector<Tensor>& H,
                                   const vector<vector<Mat>>& rep, const TensorTree& Chi, const Tree& t) {
  vector<vector<Mat>> con(t.nodes.size());
  for (int a = (int)t.nodes.size() - 1; a >= 0; --a) {
    const Node& node = t.nodes[a];
    if (node.isTop()) continue;
    const Node& p = t.nodes[node.parent];
    const Tensor& B = H[p.address];
    const size_t k = node.childIdx, n = node.shape.d.back();
    con[a].assign(H[a].s.d.back(), Mat(n, n));
    for (size_t L = 0; L < B.s.total; ++L) {
      vector<size_t> ls = indexMapping(L, B.s);
      Tensor A = Chi[p.address];
      for (size_t j = 0; j < p.children.size(); ++j) {
        if (j == k) continue;
        A = matrixTensor(rep[p.children[j]][ls[j]], A, j);
      }
      for (auto& x : A.c) x *= B[L];
      if (!p.isTop()) A = matrixTensor(con[p.address][ls.back()], A, A.s.lastIdx());
      Mat hk = contraction(Psi[p.address], A, k);
      for (size_t i = 0; i < hk.c.size(); ++i) con[a][ls[k]].c[i] += hk.c[i];
    }
  }
  return con;
}

// apply(Phi, H, rep, con, node) (TTOcontraction_Implementation.h:7-30)
Tensor ttoApply(const Tensor& Phi, const vector<Tensor>& H, const vector<vector<Mat>>& rep,
                const vector<vector<Mat>>& con, const Tree& t, const Node& node) {
  Tensor HPhi(Phi.s);
  const Tensor& B = H[node.address];
  const double eps = 1e-12;
  if (node.bottom) {
    for (size_t l = 0; l < B.s.d.back(); ++l) {
      Tensor A = matrixTensor(leafColumn(B, l, node.leaf.dim), Phi, 0);
      if (!node.isTop()) A = matrixTensor(con[node.address][l], A, A.s.lastIdx());
      for (size_t i = 0; i < A.c.size(); ++i) HPhi[i] += A[i];
    }
  } else {
    for (size_t L = 0; L < B.s.total; ++L) {
      if (std::abs(B[L]) < eps) continue;
      vector<size_t> ls = indexMapping(L, B.s);
      Tensor A = Phi;
      for (size_t k = 0; k < node.children.size(); ++k) A = matrixTensor(rep[node.children[k]][ls[k]], A, k);
      for (auto& x : A.c) x *= B[L];
      if (!node.isTop()) A = matrixTensor(con[node.address][ls.back()], A, A.s.lastIdx());
      for (size_t i = 0; i < A.c.size(); ++i) HPhi[i] += A[i];
    }
  }
  return HPhi;
}

// LayerInterface::Error (LayerInterface.cpp:6-31)
double LayerError(const Tensor& Phi, const Tensor& Chi, const HamRep& r, const Node& node) {
  double Delta = 0., norm = 1.;
  Tensor diff = Phi;
  for (size_t i = 0; i < diff.c.size(); ++i) diff[i] -= Chi[i];
  if (node.isTop()) {
    for (auto& x : diff.c) Delta += std::real(std::conj(x) * x);
    Mat S = dotProduct(Phi, Phi);
    cd tr = 0.;
    for (size_t i = 0; i < S.d1; ++i) tr += S(i, i);
    norm = std::abs(tr);
  } else {
    const Mat& rho = r.rho[node.address];
    const size_t nst = Phi.s.lastDim(), lb = Phi.s.lastBefore();
    for (size_t n = 0; n < nst; ++n)
      for (size_t m = 0; m < nst; ++m)
        for (size_t i = 0; i < lb; ++i) Delta += std::real(std::conj(diff(i, n)) * rho(n, m) * diff(i, m));
    cd tr = 0.;
    for (size_t i = 0; i < rho.d1; ++i) tr += rho(i, i);
    norm = std::abs(tr);
  }
  return std::sqrt(std::abs(Delta / norm));
}

// orthogonal (TensorTree_Implementation.h:143-171)
void orthogonal(TensorTree& P, const Tree& t) {
  for (auto& n : t.nodes) {
    if (n.isTop()) continue;
    Tensor& Phi = P[n.address];
    Mat S = dotProduct(Phi, Phi);
    Mat U;
    vector<double> ev;
    heev(S, U, ev);
    Mat SW(S.d1, S.d2);
    for (size_t j = 0; j < S.d1; ++j)
      for (size_t k = 0; k < S.d1; ++k)
        for (size_t l = 0; l < S.d1; ++l)
          SW(j, k) += U(k, l) * std::sqrt(std::max(ev[l], 0.)) * std::conj(U(j, l));
    P[n.parent] = matrixTensor(SW, P[n.parent], n.childIdx);
    gramSchmidt(Phi);
  }
}

// ---------------------------------------------------------------------------
// Bulirsch-Stoer (BS_integrator.h:31-256), T = Tensor, U = complex<double>
// ---------------------------------------------------------------------------
struct BS {
  int dim = 0, maxuse = 7, mx = 11;
  double shrink = 0.95, grow = 1.2;
  vector<int> seq;
  vector<Tensor> yvec, ytab;
  vector<double> xtab;
  size_t nDeriv = 0;
  void init(const Shape& s) {
    dim = (int)s.total;
    seq.assign(mx, 0);
    seq[0] = 2;
    seq[1] = 4;
    seq[2] = 6;
    for (int i = 3; i < mx; ++i) seq[i] = seq[i - 2] * 2;
    yvec.assign(5, Tensor(s));
    ytab.assign(maxuse, Tensor(s));
    xtab.assign(mx, 0.);
  }
  using DDX = std::function<void(double, Tensor&, const Tensor&)>;
  using ERR = std::function<double(const Tensor&, const Tensor&)>;

  void integrate(Tensor& y, double& x, double xend, double& dx, double eps, const DDX& ddx,
                 const ERR& err) {
    while (xend - x > dx * 1e-6) {
      if (x + dx > xend) dx = xend - x + 1e-10;
      y = bsstep(y, x, dx, eps, ddx, err);
    }
  }

  void clearmemory() {
    for (auto& v : yvec) v.zero();
    for (auto& v : ytab) v.zero();
    std::fill(xtab.begin(), xtab.end(), 0.);
  }

  Tensor bsstep(Tensor y, double& x, double& dx, double eps, const DDX& ddx, const ERR& err) {
    clearmemory();
    yvec[0] = y;
    yvec[3] = y;
    ddx(x, yvec[4], y);
    ++nDeriv;
    double h = dx;
    bool end = false;
    while (!end) {
      int i = 0;
      while (!(end || i >= mx)) {
        mmid(x, h / (1.0 * seq[i]), seq[i], ddx);
        const double xest = std::pow(h / (1. * seq[i]), 2);
        rzextr(i, xest, yvec[0], yvec[1], yvec[2]);
        yvec[0] = yvec[1];
        for (int j = 0; j < dim; ++j) yvec[1][j] += yvec[2][j];
        const double error = err(yvec[0], yvec[1]);
        if (error < eps) {
          x += h;
          if (i == maxuse - 1) dx = h * shrink;
          else if (i == maxuse - 2) dx = h * grow;
          else dx = (h * seq[maxuse - 2]) / (1. * seq[i]);
          end = true;
        }
        ++i;
      }
      if (!end) {
        h = h / 4.;
        for (int k = 0; k < (mx - maxuse) / 2.; ++k) h /= 2.;
        if (dx + h == dx) throw std::runtime_error("BS integrator: step size underflow");
      }
    }
    return yvec[0];
  }

  void mmid(double xstart, double h, int step, const DDX& ddx) {
    Tensor &y0 = yvec[0], &y1 = yvec[1], &y2 = yvec[2], &y3 = yvec[3], &y4 = yvec[4];
    y0 = y3;
    for (int i = 0; i < dim; ++i) y1[i] = y3[i] + h * y4[i];
    double x = xstart + h;
    ddx(x, y2, y1);
    ++nDeriv;
    for (int n = 1; n < step; ++n) {
      for (int i = 0; i < dim; ++i) y2[i] = y0[i] + 2. * h * y2[i];
      std::swap(y0.c, y1.c);  // y[0] = y[1]; y[1] = y[2] (values)
      y1 = y2;
      x += h;
      ddx(x, y2, y1);
      ++nDeriv;
    }
    for (int i = 0; i < dim; ++i) y0[i] = (y0[i] + y1[i] + h * y2[i]) / 2.;
  }

  void rzextr(int iest, double xest, Tensor& yest, Tensor& ysav, Tensor& dy) {
    vector<double> fx(maxuse, 0.);
    xtab[iest] = xest;
    if (iest == 0) {
      ysav = yest;
      dy = yest;
      ytab[0] = yest;
      return;
    }
    const int mi = iest < maxuse ? iest + 1 : maxuse;
    for (int k = 1; k < mi; ++k) fx[k] = xtab[iest - k] / xest;
    for (int j = 0; j < dim; ++j) {
      cd ddy = 0.;
      cd yy = yest[j];
      cd v = ytab[0][j];
      cd c = yy;
      ytab[0][j] = yy;
      for (int k = 1; k < mi; ++k) {
        cd bi = fx[k] * v;
        cd b = bi - c;
        if (std::abs(b) != 0.) {
          b = (c - v) / b;
          ddy = c * b;
          c = bi * b;
        } else {
          ddy = v;
        }
        v = ytab[k][j];
        ytab[k][j] = ddy;
        yy = yy + ddy;
      }
      dy[j] = ddy;
      ysav[j] = yy;
    }
  }
};

// ---------------------------------------------------------------------------
// CMF integrator (CMFIntegrator.cpp:11-249)
// ---------------------------------------------------------------------------
struct CMFResult {
  vector<double> times, energies;
  vector<cd> autocorr;
  size_t steps = 0, rejected = 0, builds = 0, derivs = 0;
  vector<double> dts;
  double dt_end = 0.;  // step size on return (written back to job.dt, CMFIntegrator.cpp:172)
};

struct CMF {
  const SOP& H;
  const Tree& t;
  cd phase;
  HamRep rep;
  vector<BS> bs;
  vector<double> dt_bs;
  double tconst = 0.25, max_increase = 2.25, eps = 1e-10;
  size_t nBuild = 0;
  CMF(const SOP& H_, const Tree& t_, cd ph) : H(H_), t(t_), phase(ph), rep(makeHamRep(H_, t_)) {
    for (auto& n : t.nodes) {
      BS b;
      b.init(n.shape);
      bs.push_back(b);
    }
  }
  void buildRep(const TensorTree& P) {
    build(rep, H, P, t, eps);
    ++nBuild;
  }
  void step(TensorTree& P, double time, double timeend, double acc_leaf) {
    for (auto& n : t.nodes) {
      const Node* np = &n;
      BS::DDX ddx = [&, np](double, Tensor& dPhi, const Tensor& Phi) {
        dPhi = LayerDerivative(H, Phi, rep, t, *np, phase);
      };
      BS::ERR er = [&, np](const Tensor& a, const Tensor& b) { return LayerError(a, b, rep, *np); };
      double lt = time;
      bs[n.address].integrate(P[n.address], lt, timeend, dt_bs[n.address], acc_leaf, ddx, er);
    }
  }
  double error(const TensorTree& A, const TensorTree& B) {
    double e = 0.;
    for (auto& n : t.nodes) e += std::pow(LayerError(A[n.address], B[n.address], rep, n), 2);
    return std::sqrt(e);
  }
  void output(double time, const TensorTree& P, const TensorTree& P0, CMFResult& res) {
    Mat hm = Expectation(H, P, rep, t);
    MatrixTree S = dotProductTree(P, P, t);
    const Mat& st = S[t.top()];
    res.times.push_back(time);
    res.energies.push_back(std::real(hm(0, 0) / st(0, 0)));
    MatrixTree ac = dotProductTree(P0, P, t);
    res.autocorr.push_back(ac[t.top()](0, 0));
  }
  CMFResult integrate(TensorTree& P, double time, double timeend, double dt, double out,
                      double acc_cmf, double acc_bs, size_t max_steps) {
    CMFResult res;
    dt_bs.assign(t.nodes.size(), dt / 2.);
    double t_next = time + out;
    bool calc_mat = true;
    TensorTree Pstart = P;
    buildRep(P);
    calc_mat = false;
    output(time, P, Pstart, res);
    double err = 0.;
    while (time + 1e-6 < timeend && res.steps < max_steps) {
      const double dtmax = std::min(timeend - time, t_next - time);
      if (dt > dtmax) dt = dtmax;
      else if (2 * dt > dtmax) dt = std::min(dt, dtmax * 0.6);
      TensorTree P0 = P;
      if (calc_mat) {
        buildRep(P);
        calc_mat = false;
      }
      const double ts = dt * tconst;
      step(P, time, time + ts, acc_bs);
      TensorTree P2 = P;
      step(P2, time + ts, time + 2 * ts, acc_bs);
      buildRep(P2);
      step(P, time + dt * tconst, time + dt * (1 - tconst), acc_bs);
      P2 = P0;
      step(P2, time, time + dt, acc_bs);
      buildRep(P2);
      step(P, time + dt * (1 - tconst), time + dt, acc_bs);
      err = error(P, P2) * 0.25;
      ++res.steps;
      if (err > acc_cmf) {
        P = P0;
        calc_mat = true;
        dt *= std::pow(acc_cmf / err, 1. / 3.) * 0.8;
        ++res.rejected;
      } else {
        time += dt;
        res.dts.push_back(dt);
        orthogonal(P, t);
        if (time + 1e-10 >= t_next) {
          output(time, P, Pstart, res);
          t_next = time + out;
        }
        dt = dt * std::min(std::pow(acc_cmf / err, 1. / 3.), max_increase);
      }
    }
    res.builds = nBuild;
    for (auto& b : bs) res.derivs += b.nDeriv;
    res.dt_end = dt;
    return res;
  }
};

// ---------------------------------------------------------------------------
// Imaginary-time eigenstates (src/Applications/Eigenstates.cpp:8-103)
// ---------------------------------------------------------------------------
// propagatorEnergies (Eigenstates.cpp:8-16): eigenvalues of S_top, ascending,
// mapped to E_i = -log(ev_i) / (2 out).
vector<double> propagatorEnergies(const TensorTree& P, const Tree& t, double out) {
  MatrixTree S = dotProductTree(P, P, t);
  Mat U;
  vector<double> ev;
  heev(S[t.top()], U, ev);
  for (auto& e : ev) e = -std::log(e) / (2. * out);
  return ev;
}

// orthonormal (TensorTree_Implementation.h:190-195): Gram-Schmidt at every node
void orthonormal(TensorTree& P, const Tree& t) {
  for (auto& n : t.nodes) gramSchmidt(P[n.address]);
}

// Eigenstate (Eigenstates.cpp:18-31): build the matrices, diagonalise the
// top-node matrix <Psi_i|H|Psi_j> and rotate the top node into its eigenbasis.
vector<double> eigenstate(TensorTree& P, const SOP& H, HamRep& rep, const Tree& t) {
  build(rep, H, P, t, 1e-10);
  Mat hm = Expectation(H, P, rep, t);
  Mat U;
  vector<double> ev;
  heev(hm, U, ev);
  const size_t top = t.top();
  P[top] = multStateArTB(U, P[top]);
  return ev;
}

struct EigenResult {
  vector<double> energies, propagator, converged;  // per output, nstates each
  size_t steps = 0, rejected = 0, builds = 0, derivs = 0;
  double dt_end = 0.;
};

// Eigenstates (Eigenstates.cpp:60-103): iterations x (CMF in imaginary time
// over [0, out - 1e-5], propagator energies, orthonormal, Eigenstate,
// overlap with the previous iterate).  converged(i) = sum_j |<Psi_i|Psi'_j>|^2
// (the s(i) of Status, Eigenstates.cpp:33-45).
EigenResult eigenstates(TensorTree& P, const SOP& H, const Tree& t, double dt, double out, size_t iterations,
                        double acc_cmf, double acc_bs, size_t max_steps) {
  EigenResult res;
  CMF cmf(H, t, cd(0., -1.));
  HamRep rep = makeHamRep(H, t);
  vector<double> e = eigenstate(P, H, rep, t);
  res.energies.insert(res.energies.end(), e.begin(), e.end());
  for (size_t it = 0; it < iterations; ++it) {
    TensorTree last = P;
    CMFResult r = cmf.integrate(P, 0., out - 1e-5, dt, out, acc_cmf, acc_bs, max_steps);
    dt = r.dt_end;
    res.steps += r.steps;
    res.rejected += r.rejected;
    res.derivs = r.derivs;  // the node integrators count across calls
    vector<double> pe = propagatorEnergies(P, t, out);
    res.propagator.insert(res.propagator.end(), pe.begin(), pe.end());
    orthonormal(P, t);
    e = eigenstate(P, H, rep, t);
    res.energies.insert(res.energies.end(), e.begin(), e.end());
    MatrixTree ov = dotProductTree(P, last, t);
    const Mat& o = ov[t.top()];
    for (size_t i = 0; i < o.d1; ++i) {
      double s = 0.;
      for (size_t j = 0; j < o.d2; ++j) s += std::norm(o(i, j));
      res.converged.push_back(s);
    }
  }
  res.builds = cmf.nBuild;
  res.dt_end = dt;
  return res;
}

// ---------------------------------------------------------------- SCF
// Krylov-sweep SCF ground-state search (SCF.cpp:36-348) on the correction
// matrices of HamiltonianRepresentation(H, tree, tree, tail = false)
// (HamiltonianRepresentation.h:19-36,90-177, HamiltonianRepresentation.cpp:
// 119-218, ActiveCounter.h:16-33).  Sparse trees are tail-cut: a term's
// tree ends at its first node with more than one active child.
struct ScfRep {
  vector<SparseMatrixTree> hMats, hCon;
  MatrixTree hCorr, hConCorr;
  vector<vector<size_t>> nAct;  // [node][term]
};

static bool mloActs(const MLO& M, size_t mode) {
  return std::find(M.modes.begin(), M.modes.end(), mode) != M.modes.end();
}

// nActives (HamiltonianRepresentation.cpp:121-138)
static size_t scfActives(const SparseMatrixTree& hm, const MLO& M, const SparseMatrixTree& hc, const Node& n) {
  size_t c = 0;
  if (n.bottom) {
    if (mloActs(M, n.leaf.mode)) ++c;
  } else {
    for (int ch : n.children)
      if (hm.active(ch)) ++c;
  }
  if (n.isTop()) return c;
  if (hc.active(n.address)) ++c;
  return c;
}

ScfRep makeScfRep(const SOP& H, const Tree& t) {
  ScfRep r;
  for (auto& M : H.terms) {
    SparseTree st = makeSparseTree(M.targetLeaves(), t, false);
    r.hMats.push_back(makeSMT(st, t));
    r.hCon.push_back(makeSMT(st, t));
  }
  for (auto& n : t.nodes) {
    const size_t d = n.shape.lastDim();
    r.hCorr.emplace_back(d, d);
    r.hConCorr.emplace_back(d, d);
    vector<size_t> na(H.terms.size());
    for (size_t l = 0; l < H.terms.size(); ++l) na[l] = scfActives(r.hMats[l], H.terms[l], r.hCon[l], n);
    r.nAct.push_back(na);
  }
  return r;
}

static size_t parentIdxOf(const Node& n) { return n.shape.lastIdx(); }

// buildUpCorrection (HamiltonianRepresentation.cpp:140-168)
static void scfUpCorrection(ScfRep& r, const SOP& H, const TensorTree& P, const Tree& t, const Node& n) {
  const size_t d = n.shape.lastDim();
  r.hCorr[n.address] = Mat(d, d);
  if (!n.bottom) {
    Tensor hA(n.shape);
    for (size_t k = 0; k < n.children.size(); ++k) {
      Tensor x = matrixTensor(r.hCorr[n.children[k]], P[n.address], k);
      for (size_t i = 0; i < hA.c.size(); ++i) hA[i] += x[i];
    }
    r.hCorr[n.address] = contraction(P[n.address], hA, parentIdxOf(n));
  }
  if (!n.isTop()) {
    for (size_t l = 0; l < H.terms.size(); ++l) {
      if (r.nAct[n.parent][l] != 1) continue;
      if (r.hMats[l].active(n.address)) {
        const Mat& h = r.hMats[l][n.address];
        for (size_t i = 0; i < h.c.size(); ++i) r.hCorr[n.address].c[i] += H.coeff[l] * h.c[i];
      }
    }
  }
}

// buildDownCorrection (HamiltonianRepresentation.cpp:170-207)
static void scfDownCorrection(ScfRep& r, const SOP& H, const TensorTree& P, const Tree& t, const Node& hole) {
  if (hole.isTop()) return;
  const Node& n = t.nodes[hole.parent];
  const Tensor& A = P[n.address];
  Tensor hA(n.shape);
  for (size_t l = 0; l < H.terms.size(); ++l) {
    if (r.nAct[n.address][l] < 2) continue;
    if (r.hMats[l].active(hole.address)) continue;
    // applyHole (SparseMatrixTreeFunctions_Implementation.h:355-374)
    Tensor tmp = A;
    for (size_t k = 0; k < n.children.size(); ++k) {
      if ((int)k == hole.childIdx || !r.hMats[l].active(n.children[k])) continue;
      tmp = matrixTensor(r.hMats[l][n.children[k]], tmp, k);
    }
    if (!n.isTop() && r.hCon[l].active(n.address)) tmp = matrixTensor(r.hCon[l][n.address], tmp, parentIdxOf(n));
    for (size_t i = 0; i < hA.c.size(); ++i) hA[i] += H.coeff[l] * tmp[i];
  }
  for (size_t k = 0; k < n.children.size(); ++k) {
    if ((int)k == hole.childIdx) continue;
    Tensor x = matrixTensor(r.hCorr[n.children[k]], A, k);
    for (size_t i = 0; i < hA.c.size(); ++i) hA[i] += x[i];
  }
  if (!n.isTop()) {
    Tensor x = matrixTensor(r.hConCorr[n.address], A, parentIdxOf(n));
    for (size_t i = 0; i < hA.c.size(); ++i) hA[i] += x[i];
  }
  r.hConCorr[hole.address] = contraction(A, hA, hole.childIdx);
}

// buildUp (HamiltonianRepresentation.h:90-105): representLayer of every
// term (SparseMatrixTreeFunctions_Implementation.h:18-52) and the up
// correction.
static void scfBuildUp(ScfRep& r, const SOP& H, const TensorTree& P, const Tree& t, const Node& n) {
  for (size_t l = 0; l < H.terms.size(); ++l) {
    SparseMatrixTree& hm = r.hMats[l];
    if (!hm.active(n.address)) continue;
    Tensor hK = n.bottom ? mloApply(H.terms[l], P[n.address], n.leaf)
                         : applyNode(hm, nullptr, P[n.address], t, n, (int)parentIdxOf(n));
    hm[n.address] = contraction(P[n.address], hK, parentIdxOf(n));
  }
  scfUpCorrection(r, H, P, t, n);
}

// buildDown (HamiltonianRepresentation.h:107-129): contractionLayer of every
// term active at the node (SparseMatrixTreeFunctions_Implementation.h:
// 145-190) and the down correction.
static void scfBuildDown(ScfRep& r, const SOP& H, const TensorTree& P, const Tree& t, const Node& n) {
  if (!n.isTop()) {
    for (size_t l = 0; l < H.terms.size(); ++l) {
      SparseMatrixTree& hc = r.hCon[l];
      if (!hc.active(n.address)) continue;
      const Node& p = t.nodes[n.parent];
      Tensor hK = applyNode(r.hMats[l], &hc, P[p.address], t, p, n.childIdx);
      hc[n.address] = contraction(P[p.address], hK, n.childIdx);
    }
  }
  scfDownCorrection(r, H, P, t, n);
}

// apply(HA, A, hMats, hCons, hCorr, hConCorr, ...) (SCF.cpp:36-81)
Tensor scfApply(const ScfRep& r, const SOP& H, const Tensor& A, const Tree& t, const Node& n) {
  Tensor HA(A.s);
  for (size_t l = 0; l < H.terms.size(); ++l) {
    Tensor hA;
    if (n.bottom) {
      if (!(r.hMats[l].active(n.address) && r.hCon[l].active(n.address))) continue;
      Tensor MA = mloApply(H.terms[l], A, n.leaf);
      hA = matrixTensor(r.hCon[l][n.address], MA, parentIdxOf(n));
    } else {
      if (r.nAct[n.address][l] <= 1) continue;
      hA = applyNode(r.hMats[l], &r.hCon[l], A, t, n, -1);
    }
    for (size_t i = 0; i < HA.c.size(); ++i) HA[i] += H.coeff[l] * hA[i];
  }
  if (!n.bottom) {
    for (size_t k = 0; k < n.children.size(); ++k) {
      Tensor x = matrixTensor(r.hCorr[n.children[k]], A, k);
      for (size_t i = 0; i < HA.c.size(); ++i) HA[i] += x[i];
    }
  }
  if (!n.isTop()) {
    Tensor x = matrixTensor(r.hConCorr[n.address], A, parentIdxOf(n));
    for (size_t i = 0; i < HA.c.size(); ++i) HA[i] += x[i];
  }
  return HA;
}

// fullContraction = trace(dotProduct(A, B)) (SCF.cpp:97-100)
static cd fullContraction(const Tensor& A, const Tensor& B) {
  Mat S = dotProduct(A, B);
  cd tr = 0.;
  for (size_t i = 0; i < S.d1; ++i) tr += S(i, i);
  return tr;
}

static double scfNormalize(Tensor& A) {  // SCF.cpp:102-107
  const double nrm = std::sqrt(std::real(fullContraction(A, A)));
  for (auto& x : A.c) x /= nrm;
  return nrm;
}

struct Krylov {
  vector<Tensor> space;
  Mat U;
  vector<double> E;
};

// solveKrylovSpace (SCF.cpp:134-196): Lanczos without re-orthogonalisation,
// k = min(nKrylov, node size) vectors; the k x k tridiagonal matrix is
// diagonalised with eigenvalues ascending and each eigenvector phased so that
// its first component is real and non-negative (the eigensolver's sign is a
// free choice: the reference's sweep is defined up to it, see scfSweep).
Krylov solveKrylov(const ScfRep& r, const SOP& H, Tensor Psi, const Tree& t, const Node& n, size_t k) {
  Krylov K;
  K.space.assign(k, Tensor(Psi.s));
  scfNormalize(Psi);
  K.space[0] = Psi;
  vector<double> alpha, beta;
  Tensor HPsi = scfApply(r, H, Psi, t, n);
  alpha.push_back(std::real(fullContraction(Psi, HPsi)));
  for (size_t i = 0; i < Psi.c.size(); ++i) Psi[i] = HPsi[i] - alpha.back() * Psi[i];
  for (size_t l = 1; l < k; ++l) {
    beta.push_back(scfNormalize(Psi));
    K.space[l] = Psi;
    HPsi = scfApply(r, H, Psi, t, n);
    alpha.push_back(std::real(fullContraction(Psi, HPsi)));
    const Tensor& v = K.space[l];
    const Tensor& vl = K.space[l - 1];
    for (size_t i = 0; i < Psi.c.size(); ++i) Psi[i] = HPsi[i] - alpha.back() * v[i] - beta.back() * vl[i];
  }
  Mat T(k, k);
  for (size_t l = 0; l < k; ++l) {
    T(l, l) = alpha[l];
    if (l == k - 1) break;
    T(l + 1, l) = beta[l];
    T(l, l + 1) = beta[l];
  }
  heev(T, K.U, K.E);
  return K;
}

// imaginaryTimePropagation (SCF.cpp:198-222).  The ground-state weight
// enters twice, as in the reference: once before the sum over k and once at
// k = 0 inside it.
static void scfItp(Tensor& Psi, const Krylov& K, double beta) {
  const size_t k = K.E.size();
  vector<cd> c(k, 0.);
  for (size_t l = 0; l < k; ++l) {
    c[l] += K.U(l, 0) * std::conj(K.U(0, 0));
    for (size_t j = 0; j < k; ++j) {
      const cd e = std::exp(-beta * (K.E[j] - K.E[0]));
      c[l] += K.U(l, j) * e * std::conj(K.U(0, j));
    }
  }
  Psi.zero();
  for (size_t l = 0; l < k; ++l)
    for (size_t i = 0; i < Psi.c.size(); ++i) Psi[i] += c[l] * K.space[l][i];
  scfNormalize(Psi);
}

// diagonalizeKrylov (SCF.cpp:224-235) over the k vectors of this solve (the
// reference sums over the nKrylov slots of the space, which equals k unless
// a node has fewer than nKrylov entries).
static void scfDiag(Tensor& Psi, const Krylov& K) {
  Psi.zero();
  for (size_t l = 0; l < K.E.size(); ++l)
    for (size_t i = 0; i < Psi.c.size(); ++i) Psi[i] += K.U(l, 0) * K.space[l][i];
}

// adjacentIndex (SCF.cpp:237-256); to < 0 stands for the null pointer
static size_t scfAdjacent(const Tree& t, const Node& from, int to) {
  if (to < 0) return parentIdxOf(from);
  for (size_t k = 0; k < from.children.size(); ++k)
    if (from.children[k] == to) return k;
  return parentIdxOf(from);
}

// qr(Tensor, mode) (Tensor_Implementation.h:864-870, toMatrix / toTensor
// :961-1000): rows (before, after) with before fastest, columns the mode.
Tensor qrTensor(const Tensor& A, size_t mode) {
  const size_t bef = A.s.bef[mode], aft = A.s.aft[mode], n = A.s.d[mode];
  Mat M(bef * aft, n);
  for (size_t b = 0; b < bef; ++b)
    for (size_t a = 0; a < n; ++a)
      for (size_t f = 0; f < aft; ++f) M(b + f * bef, a) = A.c[b + bef * (a + n * f)];
  Mat Q = householderQ(M);
  Tensor R(A.s);
  for (size_t b = 0; b < bef; ++b)
    for (size_t a = 0; a < n; ++a)
      for (size_t f = 0; f < aft; ++f) R.c[b + bef * (a + n * f)] = Q(b + f * bef, a);
  return R;
}

// scf_sweep / addNodes (SCF.cpp:14-34): the visiting order of one iteration
vector<int> scfSweepOrder(const Tree& t) {
  vector<int> s;
  std::function<void(int)> add = [&](int a) {
    s.push_back(a);
    const Node& n = t.nodes[a];
    if (n.bottom) return;
    for (int c : n.children) {
      add(c);
      s.push_back(a);
    }
  };
  add(t.top());
  return s;
}

// scf() (SCF.cpp:270-348).  energies receives the lowest Ritz value of
// every Krylov solve (what the reference prints, before its unit
// conversion).  The sweep is fixed up to the sign of each Krylov
// ground-state eigenvector; with the phase convention of solveKrylov it is
// deterministic.  With probe >= 0 the sweep stops before the Krylov solve of
// microstep `probe` and returns its node address, the state then and
// (in heff, N x N column-major) the dense effective Hamiltonian there.
int scfSweep(const SOP& H, TensorTree& P, const Tree& t, size_t nIter, size_t nKrylov, size_t nITP, double beta,
             vector<double>& energies, long probe, vector<cd>* heff) {
  const vector<int> sweep = scfSweepOrder(t);
  ScfRep r = makeScfRep(H, t);
  for (auto& n : t.nodes)
    if (!n.isTop()) scfBuildUp(r, H, P, t, n);
  long step = 0;
  for (size_t it = 0; it < nIter; ++it) {
    for (size_t l = 0; l < sweep.size(); ++l) {
      const Node& n = t.nodes[sweep[l]];
      if (step++ == probe) {
        const size_t N = n.shape.total;
        heff->assign(N * N, 0.);
        for (size_t j = 0; j < N; ++j) {
          Tensor e(n.shape);
          e[j] = 1.;
          Tensor h = scfApply(r, H, e, t, n);
          for (size_t i = 0; i < N; ++i) (*heff)[i + N * j] = h[i];
        }
        return n.address;
      }
      const size_t k = std::min(nKrylov, n.shape.total);
      Krylov K = solveKrylov(r, H, P[n.address], t, n, k);
      energies.push_back(K.E[0]);
      if (it < nITP) scfItp(P[n.address], K, beta);
      else scfDiag(P[n.address], K);
      const int next = l + 1 < sweep.size() ? sweep[l + 1] : -1;
      const size_t outIdx = scfAdjacent(t, n, next);
      Tensor PsiW = P[n.address];
      P[n.address] = qrTensor(P[n.address], outIdx);
      if (next < 0) break;
      const Node& nx = t.nodes[next];
      Mat rm = contraction(P[n.address], PsiW, outIdx);
      const size_t inIdx = scfAdjacent(t, nx, n.address);
      P[next] = matrixTensor(rm, P[next], inIdx);
      // buildSCF(H, Psi, node, next) (HamiltonianRepresentation.h:131-177)
      const bool up = !n.isTop() ? (n.parent == next) : !(nx.parent == n.address);
      if (up) scfBuildUp(r, H, P, t, n);
      else scfBuildDown(r, H, P, t, nx);
    }
  }
  return -1;
}

// ------------------------------------------- SOP application fitted into the tree
// applyOperatorIteration (ApplyFirstOrder.cpp:43-147): H|Psi> is fitted into
// a tree of Psi's shape by alternating sweeps, with fidelity() of
// TreeApplyOperator.cpp:22-35 and applyTop (:130-143).  Both files sum the
// bare products sop(i) of the SOP: its coefficients are not applied.  The
// start is TensorTree(gen, tree, false) from mt19937(seed) (the reference
// seeds it from the clock, ApplyFirstOrder.cpp:57).

// representLayer with Bra != Ket (SparseMatrixTreeFunctions_Implementation.h:18-52)
static void representLayerBK(SparseMatrixTree& h, const MLO& M, const Tensor& Bra, const Tensor& Ket, const Tree& t,
                             const Node& n) {
  if (!h.active(n.address)) return;
  Tensor hK = n.bottom ? mloApply(M, Ket, n.leaf) : applyNode(h, nullptr, Ket, t, n, (int)parentIdxOf(n));
  h[n.address] = contraction(Bra, hK, parentIdxOf(n));
}

static void representBK(SparseMatrixTree& h, const MLO& M, const TensorTree& Bra, const TensorTree& Ket,
                        const Tree& t) {
  for (int a : h.st.nodes)
    if (!t.nodes[a].isTop()) representLayerBK(h, M, Bra[a], Ket[a], t, t.nodes[a]);
}

// apply() with a density fallback for inactive holes
// (SparseMatrixTreeFunctions_Implementation.h:299-352)
static Tensor applyNodeRho(const SparseMatrixTree& mat, const SparseMatrixTree* holes, const MatrixTree* rho,
                           Tensor Phi, const Tree& t, const Node& node, int skip) {
  Tensor cur = Phi;
  if (!node.bottom) {
    for (size_t k = 0; k < node.children.size(); ++k) {
      int c = node.children[k];
      if ((int)k == skip || !mat.active(c)) continue;
      cur = matrixTensor(mat[c], cur, k);
    }
  }
  if (!node.isTop() && skip != (int)parentIdxOf(node) && holes != nullptr) {
    if (mat.st.active(node.address) && holes->active(node.address))
      cur = matrixTensor((*holes)[node.address], cur, node.shape.lastIdx());
    else if (rho != nullptr)
      cur = matrixTensor((*rho)[node.address], cur, node.shape.lastIdx());
  }
  return cur;
}

// contraction(holes, Bra, Ket, mats, &rho, stree, tree)
// (SparseMatrixTreeFunctions_Implementation.h:145-190)
static void holesRho(SparseMatrixTree& holes, const SparseMatrixTree& mats, const TensorTree& Bra,
                     const TensorTree& Ket, const MatrixTree& rho, const Tree& t) {
  for (int i = (int)holes.st.nodes.size() - 1; i >= 0; --i) {
    const int a = holes.st.nodes[i];
    const Node& n = t.nodes[a];
    if (n.isTop()) {
      holes[a] = identity(n.shape.lastDim());
      continue;
    }
    const Node& p = t.nodes[n.parent];
    Tensor hK = applyNodeRho(mats, &holes, &rho, Ket[p.address], t, p, n.childIdx);
    holes[a] = contraction(Bra[p.address], hK, n.childIdx);
  }
}

// apply(mats, Phi, M, node) (SparseMatrixTreeFunctions_Implementation.h:394-404)
static Tensor applyLocal(const SparseMatrixTree& mats, const Tensor& Phi, const MLO& M, const Tree& t,
                         const Node& n) {
  if (!mats.active(n.address)) return Phi;
  if (n.bottom) return mloApply(M, Phi, n.leaf);
  return applyNode(mats, nullptr, Phi, t, n, (int)parentIdxOf(n));
}

// applyTop (TreeApplyOperator.cpp:130-143)
static Tensor applyTopFit(const Tensor& A, const vector<SparseMatrixTree>& H, const SOP& sop, const Tree& t,
                          const Node& n) {
  Tensor out(A.s);
  for (size_t l = 0; l < sop.terms.size(); ++l) {
    Tensor x = applyLocal(H[l], A, sop.terms[l], t, n);
    for (size_t i = 0; i < out.c.size(); ++i) out[i] += x[i];
  }
  return out;
}

// applyOperatorLocal (ApplyFirstOrder.cpp:120-140)
static Tensor applyOperatorLocal(const Tensor& Phi, const MatrixTree& rho, const vector<SparseMatrixTree>& mats,
                                 const vector<SparseMatrixTree>& holes, const SOP& sop, const Tree& t,
                                 const Node& n) {
  Tensor HPhi(Phi.s);
  for (size_t l = 0; l < mats.size(); ++l) {
    Tensor m = applyLocal(mats[l], Phi, sop.terms[l], t, n);
    if (holes[l].active(n.address)) m = matrixTensor(holes[l][n.address], m, parentIdxOf(n));
    else if (!n.isTop()) m = matrixTensor(rho[n.address], m, parentIdxOf(n));
    for (size_t i = 0; i < HPhi.c.size(); ++i) HPhi[i] += m[i];
  }
  return qrTensor(HPhi, HPhi.s.lastIdx());
}

static vector<size_t> sopLeaves(const SOP& sop) {
  vector<size_t> r;
  for (auto& M : sop.terms)
    for (size_t k : M.targetLeaves())
      if (std::find(r.begin(), r.end(), k) == r.end()) r.push_back(k);
  return r;
}

// fidelity(Psi, Psilast, u, tree) (TreeApplyOperator.cpp:22-35)
double fitFidelity(const TensorTree& HPsi, const TensorTree& Psi, const SOP& sop, const Tree& t) {
  SparseTree st = makeSparseTree(sopLeaves(sop), t, true);
  vector<SparseMatrixTree> um;
  for (size_t l = 0; l < sop.terms.size(); ++l) {
    um.push_back(makeSMT(st, t));
    representBK(um[l], sop.terms[l], HPsi, Psi, t);
  }
  const Node& top = t.nodes[t.top()];
  Tensor U = applyTopFit(Psi[top.address], um, sop, t, top);
  return std::norm(dotProduct(HPsi[top.address], U)(0, 0));
}

// applyOperatorIteration(Psi, rho, sop, tree) (ApplyFirstOrder.cpp:43-79) and
// its sweep (:81-118).  fids receives the fidelity after every iteration.
void applyOperatorFit(const SOP& sop, TensorTree& Psi, const Tree& t, uint64_t seed, size_t nIter,
                      vector<double>& fids) {
  if (sop.terms.size() == 1) {  // applyReference on the bottom tensors
    for (auto& n : t.nodes)
      if (n.bottom) Psi[n.address] = mloApply(sop.terms[0], Psi[n.address], n.leaf);
    fids.push_back(1.0);
    return;
  }
  SparseTree st = makeSparseTree(sopLeaves(sop), t, false);
  MatrixTree rho = rhoTree(Psi, t);
  std::mt19937 gen((unsigned)seed);
  TensorTree HPsi = makeTensorTree(t);
  fillRandom(HPsi, t, gen, false);
  for (auto& n : t.nodes)
    if (!st.active(n.address)) HPsi[n.address] = Psi[n.address];
  double f = 0., f_last = 1.;
  for (size_t it = 0; it < nIter; ++it) {
    vector<SparseMatrixTree> Hm, Hh;
    for (size_t l = 0; l < sop.terms.size(); ++l) {
      Hm.push_back(makeSMT(st, t));
      representBK(Hm[l], sop.terms[l], HPsi, Psi, t);
    }
    for (size_t l = 0; l < sop.terms.size(); ++l) {
      Hh.push_back(makeSMT(st, t));
      holesRho(Hh[l], Hm[l], HPsi, Psi, rho, t);
    }
    for (int a : st.nodes) {
      const Node& n = t.nodes[a];
      const bool top = n.isTop() || !st.active(n.parent);
      if (!top) {
        HPsi[a] = applyOperatorLocal(Psi[a], rho, Hm, Hh, sop, t, n);
        for (size_t l = 0; l < sop.terms.size(); ++l) representLayerBK(Hm[l], sop.terms[l], HPsi[a], Psi[a], t, n);
      } else {
        HPsi[a] = applyTopFit(Psi[a], Hm, sop, t, n);
        gramSchmidt(HPsi[a]);
      }
    }
    f_last = f;
    f = fitFidelity(HPsi, Psi, sop, t);
    fids.push_back(f);
    if ((1. - f) < 1e-6 || (f - f_last) < 1e-10) break;
  }
  Psi = HPsi;
}

// ---------------------------------------------------------------------------
// Model = tree + SOP (+ state), the unit the C-ABI hands around.
// ---------------------------------------------------------------------------
struct Model {
  Tree tree;
  SOP H;
  HamRep rep;
  bool built = false;
  size_t psiSize() const {
    size_t s = 0;
    for (auto& n : tree.nodes) s += n.shape.total;
    return s;
  }
  TensorTree unpack(const double* p) const {
    TensorTree P = makeTensorTree(tree);
    size_t o = 0;
    for (auto& T : P)
      for (auto& x : T.c) {
        x = cd(p[2 * o], p[2 * o + 1]);
        ++o;
      }
    return P;
  }
  void pack(const TensorTree& P, double* p) const {
    size_t o = 0;
    for (auto& T : P)
      for (auto& x : T.c) {
        p[2 * o] = x.real();
        p[2 * o + 1] = x.imag();
        ++o;
      }
  }
};

}  // namespace orc

// ===========================================================================
// C-ABI used by the pytest suite (ctypes) and bench.py's CPU baseline.
// Complex data is interleaved (re, im) double.
// ===========================================================================
using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// ---- node level -----------------------------------------------------------
void orc_matvec(double* C, const double* B, const double* M, long a, long b, long c, int add) {
  matvec((cd*)C, (const cd*)B, (const cd*)M, a, b, c, add != 0);
}
void orc_tmatvec(double* C, const double* B, const double* M, long a, long b, long c, int add) {
  tmatvec((cd*)C, (const cd*)B, (const cd*)M, a, b, c, add != 0);
}
void orc_ctmatvec(double* C, const double* B, const double* M, long a, long b, long c, int add) {
  ctmatvec((cd*)C, (const cd*)B, (const cd*)M, a, b, c, add != 0);
}
void orc_rhomat(const double* bra, const double* ket, double* M, long a, long b, long c) {
  rhomat((const cd*)bra, (const cd*)ket, (cd*)M, a, b, c);
}
void orc_mode_product(double* C, const double* M, const double* B, long before, long nB, long nC,
                      long after, int add) {
  matrixTensorRect((cd*)C, (const cd*)M, (const cd*)B, before, nB, nC, after, add != 0);
}
void orc_hole_contraction(double* S, const double* A, const double* B, long before, long nA, long nB,
                          long after, int accumulate) {
  contractionRect((cd*)S, (const cd*)A, (const cd*)B, before, nA, nB, after, accumulate != 0);
}
// Batch of independent (matrixTensor, contraction) pairs on order-3 tensors:
// the kernel microbenchmark workload.  shapes: 3 ints per tensor, modes: 1 int.
// Offsets are in complex elements.  C has the shape of A.
void orc_microbench_batch(long count, const long* shapes, const long* modes, const long* offT,
                          const long* offM, const double* A, const double* B, const double* M,
                          double* C, double* S) {
  for (long t = 0; t < count; ++t) {
    const long d0 = shapes[3 * t], d1 = shapes[3 * t + 1], d2 = shapes[3 * t + 2];
    const long k = modes[t];
    const long dims[3] = {d0, d1, d2};
    long before = 1, after = 1;
    for (long i = 0; i < k; ++i) before *= dims[i];
    for (long i = k + 1; i < 3; ++i) after *= dims[i];
    const long n = dims[k];
    const cd* a = (const cd*)A + offT[t];
    const cd* b = (const cd*)B + offT[t];
    const cd* m = (const cd*)M + offM[t];
    matvec((cd*)C + offT[t], a, m, n, before, after, false);
    rhomat(a, b, (cd*)S + offM[t], n, before, after);
  }
}
void orc_gram_schmidt(double* A, long before, long n) {
  Tensor T(Shape({(size_t)before, (size_t)n}));
  std::memcpy(T.c.data(), A, sizeof(cd) * before * n);
  gramSchmidt(T);
  std::memcpy(A, T.c.data(), sizeof(cd) * before * n);
}
void orc_heev(const double* A, long n, double* U, double* ev) {
  Mat M(n, n);
  std::memcpy(M.c.data(), A, sizeof(cd) * n * n);
  Mat V;
  vector<double> e;
  heev(M, V, e);
  std::memcpy(U, V.c.data(), sizeof(cd) * n * n);
  for (long i = 0; i < n; ++i) ev[i] = e[i];
}
void orc_rmatvec(double* C, const double* B, const double* M, long a, long b, long c, int add) {
  rmatvec(C, B, M, a, b, c, add != 0);
}
void orc_trmatvec(double* C, const double* B, const double* M, long a, long b, long c, int add) {
  trmatvec(C, B, M, a, b, c, add != 0);
}
void orc_inverse(const double* A, long n, double* R) {
  Mat M(n, n);
  std::memcpy(M.c.data(), A, sizeof(cd) * n * n);
  Mat X = inverse(M);
  std::memcpy(R, X.c.data(), sizeof(cd) * n * n);
}
void orc_qr(const double* A, long m, long n, double* Q) {
  Mat M(m, n);
  std::memcpy(M.c.data(), A, sizeof(cd) * m * n);
  Mat X = householderQ(M);
  std::memcpy(Q, X.c.data(), sizeof(cd) * m * n);
}
void orc_reg_inverse(const double* A, long n, double eps, double* R) {
  Mat M(n, n);
  std::memcpy(M.c.data(), A, sizeof(cd) * n * n);
  Mat X = regularizedInverse(M, eps);
  std::memcpy(R, X.c.data(), sizeof(cd) * n * n);
}
void orc_project_out(const double* A, const double* B, long before, long n, double* out) {
  Tensor a(Shape({(size_t)before, (size_t)n})), b(Shape({(size_t)before, (size_t)n}));
  std::memcpy(a.c.data(), A, sizeof(cd) * before * n);
  std::memcpy(b.c.data(), B, sizeof(cd) * before * n);
  Tensor r = projectOut(a, b);
  std::memcpy(out, r.c.data(), sizeof(cd) * before * n);
}

// ---- trees ------------------------------------------------------------------
// Writes the compact text of TreeFactory::balancedTree into buf (returns length
// needed including NUL).
long orc_balanced_tree_text(long num_leaves, long dim_leaves, long dim_nodes, long dim_inc,
                            double omega, double r0, double wfr0, double wfomega, char* buf,
                            long buflen) {
  std::string s = balancedTreeText(num_leaves, dim_leaves, dim_nodes, dim_inc, omega, r0, wfr0, wfomega);
  if (buf && buflen > (long)s.size()) std::memcpy(buf, s.c_str(), s.size() + 1);
  return (long)s.size() + 1;
}

void* orc_model_create(const char* tree_text) {
  try {
    auto* m = new Model();
    m->tree = readTree(tree_text);
    return m;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_model_destroy(void* h) { delete (Model*)h; }
long orc_model_nnodes(void* h) { return (long)((Model*)h)->tree.nodes.size(); }
long orc_model_psi_size(void* h) { return (long)((Model*)h)->psiSize(); }
// info: [parent, childIdx, bottom, layer, order, nchildren, mode(-1), dims...(order)]
long orc_model_node_info(void* h, long addr, long* out, long outlen) {
  const Node& n = ((Model*)h)->tree.nodes.at(addr);
  vector<long> v = {n.parent, n.childIdx, n.bottom ? 1 : 0, n.layer, (long)n.shape.order(),
                    (long)n.children.size(), n.bottom ? (long)n.leaf.mode : -1};
  for (auto d : n.shape.d) v.push_back((long)d);
  for (int c : n.children) v.push_back(c);
  if (out && outlen >= (long)v.size()) std::copy(v.begin(), v.end(), out);
  return (long)v.size();
}
// leaf basis data for mode: x (dim), kin (dim*dim real, column-major), trafo
void orc_model_leaf(void* h, long mode, double* x, double* kin, double* trafo) {
  Model* m = (Model*)h;
  const Leaf& L = m->tree.nodes[m->tree.leafNode.at(mode)].leaf;
  std::copy(L.x.begin(), L.x.end(), x);
  std::copy(L.kin.begin(), L.kin.end(), kin);
  std::copy(L.trafo.begin(), L.trafo.end(), trafo);
}

// Hamiltonian: term coefficients (re,im), factor lists with kinds; kMatrix
// factors index into a pool of interleaved complex d x d matrices (offsets in
// complex elements).
int orc_model_set_hamiltonian(void* h, long nterms, const double* coeff, const long* term_off,
                              const long* fmode, const long* fkind, const long* fmat_off,
                              const double* pool) {
  Model* m = (Model*)h;
  try {
    m->H = SOP();
    for (long l = 0; l < nterms; ++l) {
      MLO M;
      for (long f = term_off[l]; f < term_off[l + 1]; ++f) {
        LeafOp op;
        op.kind = (int)fkind[f];
        const size_t mode = (size_t)fmode[f];
        if (mode >= m->tree.leafNode.size()) throw std::runtime_error("factor mode out of range");
        const Leaf& L = m->tree.nodes[m->tree.leafNode[mode]].leaf;
        const size_t d = L.dim;
        if (op.kind == kMatrix) {
          op.m = Mat(d, d);
          std::memcpy(op.m.c.data(), pool + 2 * fmat_off[f], sizeof(cd) * d * d);
        } else if (op.kind == kPoly) {
          // a_T T + diag(a_1 x + a_2 x^2 + a_4 x^4) on this leaf's DVR grid
          const double* a = pool + 2 * fmat_off[f];
          op.m = Mat(d, d);
          for (size_t j = 0; j < d; ++j)
            for (size_t i = 0; i < d; ++i) op.m(i, j) = a[0] * L.kin[j * d + i];
          for (size_t i = 0; i < d; ++i) {
            const double x = L.x[i];
            op.m(i, i) += a[1] * x + a[2] * x * x + a[3] * x * x * x * x;
          }
        } else if (op.kind < 0 || op.kind > kPoly) {
          throw std::runtime_error("unknown leaf operator kind");
        }
        M.ops.push_back(op);
        M.modes.push_back(mode);
      }
      m->H.terms.push_back(M);
      m->H.coeff.push_back(cd(coeff[2 * l], coeff[2 * l + 1]));
    }
    m->rep = makeHamRep(m->H, m->tree);
    m->built = false;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

void orc_fill_random(void* h, unsigned long seed, int delta_lowest, double* psi) {
  Model* m = (Model*)h;
  std::mt19937 gen((std::mt19937::result_type)seed);
  TensorTree P = makeTensorTree(m->tree);
  fillRandom(P, m->tree, gen, delta_lowest != 0);
  m->pack(P, psi);
}

void orc_orthogonal(void* h, double* psi) {
  Model* m = (Model*)h;
  TensorTree P = m->unpack(psi);
  orthogonal(P, m->tree);
  m->pack(P, psi);
}

// Sparse tree of term l: writes active addresses, returns count.
long orc_sparse_tree(void* h, long l, long* out) {
  Model* m = (Model*)h;
  const auto& st = m->rep.hMats.at(l).st;
  if (out)
    for (size_t i = 0; i < st.nodes.size(); ++i) out[i] = st.nodes[i];
  return (long)st.nodes.size();
}
long orc_sparse_tree_modes(void* h, const long* modes, long nmodes, int tail, long* out) {
  Model* m = (Model*)h;
  vector<size_t> md(modes, modes + nmodes);
  SparseTree st = makeSparseTree(md, m->tree, tail != 0);
  if (out)
    for (size_t i = 0; i < st.nodes.size(); ++i) out[i] = st.nodes[i];
  return (long)st.nodes.size();
}

int orc_build(void* h, const double* psi, double eps) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    build(m->rep, m->H, P, m->tree, eps);
    m->built = true;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
// which: 0 = rho, 1 = rhoInv (addr ignored for l), 2 = hMat[l][addr], 3 = hole[l][addr]
int orc_get_matrix(void* h, int which, long l, long addr, double* out) {
  Model* m = (Model*)h;
  const Mat* M = nullptr;
  if (which == 0) M = &m->rep.rho.at(addr);
  else if (which == 1) M = &m->rep.rhoInv.at(addr);
  else if (which == 2) {
    if (!m->rep.hMats.at(l).active(addr)) return 1;
    M = &m->rep.hMats[l][addr];
  } else {
    if (!m->rep.hCon.at(l).active(addr)) return 1;
    M = &m->rep.hCon[l][addr];
  }
  std::memcpy(out, M->c.data(), sizeof(cd) * M->c.size());
  return 0;
}

// Derivative (HamiltonianRepresentation.cpp:74-81) for the whole tree:
// build + LayerDerivative at every node.
// Threads the build and the per-node loop use (1 unless built with OpenMP).
int orc_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_derivative(void* h, const double* psi, double* dpsi, double phase_re, double phase_im,
                   double eps) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    build(m->rep, m->H, P, m->tree, eps);
    m->built = true;
    TensorTree D = makeTensorTree(m->tree);
#pragma omp parallel for schedule(dynamic)
    for (long a = 0; a < (long)m->tree.nodes.size(); ++a) {
      const Node& n = m->tree.nodes[a];
      D[n.address] = LayerDerivative(m->H, P[n.address], m->rep, m->tree, n, cd(phase_re, phase_im));
    }
    m->pack(D, dpsi);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
// Energy gradient dE/dA* at every node (orthonormal lower nodes): build, then
// EnergyGradient.  *energy receives E.
int orc_gradient(void* h, const double* psi, double* grad, double eps, double* energy) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    build(m->rep, m->H, P, m->tree, eps);
    m->built = true;
    double E = 0.0;
    TensorTree G = EnergyGradient(m->H, P, m->rep, m->tree, E);
    m->pack(G, grad);
    if (energy) *energy = E;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
// TTNO application at every node: representation (bra, ket), contraction,
// then apply() to the ket's node tensors.  bond[a] = D_a (1 at the top);
// B holds the operator tensors in node-address order, shapes as above.
// expect receives rep[top][0] = <bra|H|ket> (top with one state).
int orc_ttno_apply(void* h, const long* bond, const double* B, const double* bra, const double* ket, double* out,
                   double* expect) {
  Model* m = (Model*)h;
  try {
    const Tree& t = m->tree;
    vector<Tensor> H;
    size_t o = 0;
    for (const Node& n : t.nodes) {
      vector<size_t> d;
      if (n.bottom) d.push_back(n.leaf.dim * n.leaf.dim);
      else
        for (int c : n.children) d.push_back((size_t)bond[c]);
      d.push_back((size_t)bond[n.address]);
      if (n.isTop() && bond[n.address] != 1) throw std::runtime_error("ttno: top bond dimension must be 1");
      Tensor T{Shape(d)};
      for (auto& x : T.c) {
        x = cd(B[2 * o], B[2 * o + 1]);
        ++o;
      }
      H.push_back(T);
    }
    TensorTree Psi = m->unpack(bra), Chi = m->unpack(ket);
    auto rep = ttoRepresentation(Psi, H, Chi, t);
    auto con = ttoContraction(Psi, H, rep, Chi, t);
    TensorTree D = makeTensorTree(t);
    for (const Node& n : t.nodes) D[n.address] = ttoApply(Chi[n.address], H, rep, con, t, n);
    m->pack(D, out);
    if (expect) {
      const Mat& e = rep[t.top()][0];
      expect[0] = e.c[0].real();
      expect[1] = e.c[0].imag();
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
// Apply at every node with the current representation (no projector).
int orc_apply_all(void* h, const double* psi, double* out) {
  Model* m = (Model*)h;
  if (!m->built) return 1;
  TensorTree P = m->unpack(psi);
  TensorTree D = makeTensorTree(m->tree);
  for (auto& n : m->tree.nodes) D[n.address] = Apply(m->H, P[n.address], m->rep, m->tree, n);
  m->pack(D, out);
  return 0;
}
// Energy = real(Expectation / S_top) with the current representation.
double orc_energy(void* h, const double* psi) {
  Model* m = (Model*)h;
  TensorTree P = m->unpack(psi);
  Mat hm = Expectation(m->H, P, m->rep, m->tree);
  MatrixTree S = dotProductTree(P, P, m->tree);
  return std::real(hm(0, 0) / S[m->tree.top()](0, 0));
}
void orc_dot_product_top(void* h, const double* bra, const double* ket, double* out) {
  Model* m = (Model*)h;
  TensorTree A = m->unpack(bra), B = m->unpack(ket);
  MatrixTree S = dotProductTree(A, B, m->tree);
  out[0] = S[m->tree.top()](0, 0).real();
  out[1] = S[m->tree.top()](0, 0).imag();
}

// CMF propagation; outputs at every 'out' time: times, energies, autocorr (re,im).
// Returns number of outputs, or -1 on error.  stats: [steps, rejected, builds, derivs]
long orc_cmf(void* h, double* psi, double t0, double tend, double dt, double out, double acc_cmf,
             double acc_bs, double phase_re, double phase_im, long max_steps, long max_out,
             double* times, double* energies, double* autocorr, long* stats) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    CMF cmf(m->H, m->tree, cd(phase_re, phase_im));
    CMFResult r = cmf.integrate(P, t0, tend, dt, out, acc_cmf, acc_bs, (size_t)max_steps);
    m->pack(P, psi);
    long n = std::min((long)r.times.size(), max_out);
    for (long i = 0; i < n; ++i) {
      times[i] = r.times[i];
      energies[i] = r.energies[i];
      autocorr[2 * i] = r.autocorr[i].real();
      autocorr[2 * i + 1] = r.autocorr[i].imag();
    }
    if (stats) {
      stats[0] = (long)r.steps;
      stats[1] = (long)r.rejected;
      stats[2] = (long)r.builds;
      stats[3] = (long)r.derivs;
    }
    return (long)r.times.size();
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// TensorTree<cd>::write (TensorTree_Implementation.h:79-92 with
// Tensor::write, Tensor_Implementation.h:204-221, and TensorShape::write,
// TensorShape.cpp:68-81) into buf; returns the byte count (buf may be null).
long orc_tensor_tree_write(void* h, const double* psi, unsigned char* buf, long cap) {
  Model* m = (Model*)h;
  TensorTree P = m->unpack(psi);
  std::string out;
  auto put = [&](const void* p, size_t n) { out.append((const char*)p, n); };
  put("TTre", 4);
  const int32_t nn = (int32_t)P.size();
  put(&nn, 4);
  for (const Tensor& T : P) {
    put("TENS", 4);
    put("TDIM", 4);
    const int32_t order = (int32_t)T.s.order();
    put(&order, 4);
    for (size_t k = 0; k < T.s.order(); ++k) {
      const int32_t d = (int32_t)T.s.d[k];
      put(&d, 4);
    }
    const int32_t size = sizeof(cd);
    put(&size, 4);
    for (const cd& x : T.c) put(&x, sizeof(cd));
  }
  if (buf && (long)out.size() <= cap) std::memcpy(buf, out.data(), out.size());
  return (long)out.size();
}

// Krylov-sweep SCF (scfSweep): psi is updated in place; energies (cap
// entries) receives the lowest Ritz value of every microstep.  Returns the
// number of microsteps, -1 on error.
long orc_scf(void* h, double* psi, long nIter, long nKrylov, long nITP, double beta, double* energies, long cap) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    vector<double> e;
    scfSweep(m->H, P, m->tree, (size_t)nIter, (size_t)nKrylov, (size_t)nITP, beta, e, -1, nullptr);
    m->pack(P, psi);
    for (long i = 0; i < (long)e.size() && i < cap; ++i) energies[i] = e[i];
    return (long)e.size();
  } catch (std::exception& ex) {
    g_err = ex.what();
    return -1;
  }
}

// Runs the SCF sweep (nITP, beta as in orc_scf) up to microstep `step` and
// stops before its Krylov
// solve: psi receives the state then, heff (N x N complex, column-major,
// cap complex entries) the dense effective Hamiltonian of the node it is at.
// Returns that node's address.
long orc_scf_probe(void* h, double* psi, long nKrylov, long nITP, double beta, long step, double* heff,
                   long cap) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    vector<double> e;
    vector<cd> H;
    const size_t iters = (size_t)(step / (long)scfSweepOrder(m->tree).size()) + 1;
    const int a = scfSweep(m->H, P, m->tree, iters, (size_t)nKrylov, (size_t)nITP, beta, e, step, &H);
    m->pack(P, psi);
    for (long i = 0; i < (long)H.size() && i < cap; ++i) {
      heff[2 * i] = H[i].real();
      heff[2 * i + 1] = H[i].imag();
    }
    return a;
  } catch (std::exception& ex) {
    g_err = ex.what();
    return -1;
  }
}

// The node addresses one SCF iteration visits (scf_sweep); returns the count.
long orc_scf_sweep_order(void* h, long* out, long cap) {
  vector<int> s = scfSweepOrder(((Model*)h)->tree);
  for (long i = 0; i < (long)s.size() && i < cap; ++i) out[i] = s[i];
  return (long)s.size();
}

// SOP application fitted into the tree (applyOperatorFit): psi in, the fit
// out; fids (cap entries) receives the fidelity of every iteration.  Returns
// the number of iterations, -1 on error.
long orc_apply_fit(void* h, double* psi, unsigned long seed, long n_iter, double* fids, long cap) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    vector<double> f;
    applyOperatorFit(m->H, P, m->tree, seed, (size_t)n_iter, f);
    m->pack(P, psi);
    for (long i = 0; i < (long)f.size() && i < cap; ++i) fids[i] = f[i];
    return (long)f.size();
  } catch (std::exception& ex) {
    g_err = ex.what();
    return -1;
  }
}

// Top-node matrices after a build from psi: H(i,j) = <Psi_i|H|Psi_j>
// (Expectation) and S = dotProduct(Psi, Psi)[top], column-major interleaved.
int orc_top_matrices(void* h, const double* psi, double* hmat, double* smat) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    build(m->rep, m->H, P, m->tree, 1e-10);
    m->built = true;
    Mat hm = Expectation(m->H, P, m->rep, m->tree);
    MatrixTree S = dotProductTree(P, P, m->tree);
    const Mat& st = S[m->tree.top()];
    for (size_t i = 0; i < hm.c.size(); ++i) {
      hmat[2 * i] = hm.c[i].real();
      hmat[2 * i + 1] = hm.c[i].imag();
      smat[2 * i] = st.c[i].real();
      smat[2 * i + 1] = st.c[i].imag();
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Imaginary-time eigenstates; psi in/out.  energies: (iterations + 1) x
// nstates, propagator and converged: iterations x nstates.  stats: [steps,
// rejected, builds, derivs].  Returns nstates, or -1 on error.
long orc_eigenstates(void* h, double* psi, double dt, double out, long iterations, double acc_cmf, double acc_bs,
                     long max_steps, double* energies, double* propagator, double* converged, long* stats,
                     double* dt_end) {
  Model* m = (Model*)h;
  try {
    TensorTree P = m->unpack(psi);
    EigenResult r = eigenstates(P, m->H, m->tree, dt, out, (size_t)iterations, acc_cmf, acc_bs, (size_t)max_steps);
    m->pack(P, psi);
    std::copy(r.energies.begin(), r.energies.end(), energies);
    std::copy(r.propagator.begin(), r.propagator.end(), propagator);
    std::copy(r.converged.begin(), r.converged.end(), converged);
    if (stats) {
      stats[0] = (long)r.steps;
      stats[1] = (long)r.rejected;
      stats[2] = (long)r.builds;
      stats[3] = (long)r.derivs;
    }
    if (dt_end) *dt_end = r.dt_end;
    return (long)m->tree.nodes[m->tree.top()].shape.lastDim();
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

}  // extern "C"
