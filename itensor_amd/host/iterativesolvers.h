// davidson (iterativesolvers.h:99-451 of the reference) over device ITensors.
//
// BigMatrixT needs product(ITensor const&, ITensor&) const and size() const
// (the reference's requirement, iterativesolvers.h:28).  The vectors, the
// Gram-Schmidt pass, the residual and its norm are FP64 device work; the
// Rayleigh-Ritz matrix (at most (maxiter+2)^2) is solved on the host in
// FP64 by cyclic Jacobi (it is real symmetric; the reference diagonalises
// the same numbers as a complex Hermitian matrix with zheev).
#pragma once

#include <cmath>
#include <vector>

#include "itensor.h"

namespace itensor {

namespace detail {
// Eigen-decomposition of a small real symmetric matrix (row-major n x n):
// eigenvalues ascending, eigenvectors in the columns of V (row-major).
void smallSymEig(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V);
}  // namespace detail

struct DavidsonStats {
  int products = 0;
  int iterations = 0;
  double qnorm = 0;
};

template <class BigMatrixT>
Real davidson(BigMatrixT const &A, ITensor &phi, Args const &args, DavidsonStats *stats = nullptr) {
  const size_t maxiter = size_t(args.getInt("MaxIter", 2));
  const double errgoal = args.getReal("ErrGoal", 1e-14);
  const size_t miniter = size_t(args.getInt("MinIter", 1));
  const double Approx0 = 1e-12;
  const uint64_t seed = uint64_t(args.getInt("RandomSeed", 12345));

  double nrm = norm(phi);
  int tries = 0;
  while (nrm == 0.0) {
    randomize(phi, seed + uint64_t(tries++));
    nrm = norm(phi);
  }
  phi *= 1.0 / nrm;
  const size_t maxsize = size_t(A.size());
  const size_t actual_maxiter = std::min(maxiter, maxsize - 1);
  std::vector<ITensor> V(actual_maxiter + 2), AV(actual_maxiter + 2);
  const size_t MS = actual_maxiter + 2;
  std::vector<double> M(MS * MS, NAN);
  double last_lambda = 1000.0, lambda = NAN, qnorm = NAN;
  int nprod = 0, iters = 0;

  V[0] = phi;
  A.product(V[0], AV[0]);
  ++nprod;
  const double initEn = dotDag(V[0], AV[0]);
  ITensor phi_t = phi;

  for (size_t ii = 0; ii <= actual_maxiter; ++ii) {
    const size_t ni = ii + 1;
    ITensor q;
    if (ii == 0) {
      lambda = initEn;
      M[0] = lambda;
      q = lincomb({&AV[0], &V[0]}, {1.0, -lambda});
    } else {
      const int m = int(ii + 1);
      std::vector<double> Mr(size_t(m * m)), w, U;
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c) Mr[size_t(r * m + c)] = M[size_t(r) * MS + size_t(c)];
      detail::smallSymEig(m, Mr, w, U);
      lambda = w[0];
      // phi_t = sum_k U_k0 V_k, q = sum_k U_k0 AV_k - lambda phi_t, both
      // negated when U_00 < 0 (iterativesolvers.h:257-283)
      const double sgn = U[0] < 0 ? -1.0 : 1.0;
      std::vector<ITensor const *> vp, qp;
      std::vector<double> vc, qc;
      for (int k = 0; k <= int(ii); ++k) {
        const double u = sgn * U[size_t(k * m + 0)];
        vp.push_back(&V[size_t(k)]);
        vc.push_back(u);
        qp.push_back(&AV[size_t(k)]);
        qc.push_back(u);
      }
      for (int k = 0; k <= int(ii); ++k) {
        qp.push_back(&V[size_t(k)]);
        qc.push_back(-lambda * vc[size_t(k)]);
      }
      phi_t = lincomb(vp, vc);
      q = lincomb(qp, qc);
    }
    qnorm = norm(q);
    const bool converged = (qnorm < errgoal && std::abs(lambda - last_lambda) < errgoal) ||
                           qnorm < std::max(Approx0, errgoal * 1e-3);
    last_lambda = lambda;
    if (qnorm < 1e-20 || (converged && ii >= miniter) || ii == actual_maxiter) break;

    // one classical Gram-Schmidt pass (iterativesolvers.h:305-357): all
    // projections from one batched reduction, one linear combination
    int Npass = 1, pass = 1, tot_pass = 0;
    bool stop = false;
    while (pass <= Npass) {
      ++tot_pass;
      std::vector<ITensor const *> vk;
      for (size_t k = 0; k < ni; ++k) vk.push_back(&V[k]);
      std::vector<double> Vq = dotDagMany(vk, q);
      std::vector<ITensor const *> gp{&q};
      std::vector<double> gc{1.0};
      for (size_t k = 0; k < ni; ++k) {
        gp.push_back(&V[k]);
        gc.push_back(-Vq[k]);
      }
      q = lincomb(gp, gc);
      double qnrm = norm(q);
      if (qnrm < 1e-10) {
        q = V[ni - 1];
        randomize(q, seed + 7919u * uint64_t(ni) + uint64_t(tot_pass));
        qnrm = norm(q);
        --pass;
        if (ni >= maxsize || tot_pass > Npass * 3) {
          stop = true;
          break;
        }
      }
      q *= 1.0 / qnrm;
      ++pass;
    }
    if (stop) break;
    V[ni] = q;
    A.product(V[ni], AV[ni]);
    ++nprod;
    std::vector<ITensor const *> vk;
    for (size_t k = 0; k <= ni; ++k) vk.push_back(&V[k]);
    const std::vector<double> z = dotDagMany(vk, AV[ni]);
    for (size_t k = 0; k <= ni; ++k) {
      M[k * MS + ni] = z[k];
      M[ni * MS + k] = z[k];
    }
    ++iters;
  }
  phi = phi_t;
  if (stats) {
    stats->products = nprod;
    stats->iterations = iters;
    stats->qnorm = qnorm;
  }
  return lambda;
}

}  // namespace itensor
