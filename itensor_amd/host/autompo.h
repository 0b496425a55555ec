// Site sets and a sum-of-terms MPO builder for the ITensor-compatible host
// layer.
//
// Site sets follow the reference's site types: SpinHalf (sites/spinhalf.h,
// basis Up, Dn; QN "Sz" in units of 2 Sz) and Electron (sites/electron.h,
// basis Emp, Up, Dn, UpDn; QNs "Nf" and "Sz").  Operator matrices are
// given as <s'|O|s>; the fermionic operators C* carry the on-site sign
// (-1)^{n_up} on the down channel, as Electron's Cdn / Cdagdn do, so that
// c_{i,sigma} = F_1 ... F_{i-1} C_{i,sigma} with F = (-1)^{n_up + n_dn}.
//
// AutoMPO here builds the MPO of a sum of one- and two-site terms exactly
// (no compression), as a finite-state machine over the MPO link: state
// "done", state "start", and one state per (first site i, first operator A)
// of the open two-site terms, alive on the links between i and the last
// second site of those terms.  Fermionic pairs get the Jordan-Wigner
// string: for i < j, A_i B_j = (A F)_i F_{i+1} ... F_{j-1} B_j.  The
// reference's AutoMPO (mps/autompo.cc:982-1496) reaches a smaller link
// dimension by SVD compression; the operator it represents is the same.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "mps.h"

namespace itensor {

// A local operator: m[s + d * sp] = <sp|O|s>, its QN flux and parity.
struct SiteOp {
  int d = 0;
  std::vector<double> m;
  QN flux;
  bool fermionic = false;
};

class SiteSet {
 public:
  virtual ~SiteSet() = default;
  int length() const { return int(s_.size()); }
  Index const &operator()(int i) const { return s_.at(size_t(i - 1)); }  // 1-based
  virtual SiteOp op(std::string const &name, int i) const = 0;
  virtual std::shared_ptr<SiteSet> clone() const = 0;

 protected:
  std::vector<Index> s_;
};

// S = 1/2 sites: Id, Sz, S+ (Sp), S- (Sm), Sx
class SpinHalf : public SiteSet {
 public:
  explicit SpinHalf(int N, bool conserveQNs = true);
  SiteOp op(std::string const &name, int i) const override;
  std::shared_ptr<SiteSet> clone() const override { return std::make_shared<SpinHalf>(*this); }
};

// Spinful fermion sites: Id, F, Fup, Fdn, Nup, Ndn, Nupdn, Ntot, Cup, Cdn,
// Cdagup, Cdagdn (fermionic), Aup, Adn, Adagup, Adagdn (no on-site sign,
// bosonic), Sz, S+, S-
class Electron : public SiteSet {
 public:
  explicit Electron(int N, bool conserveQNs = true);
  SiteOp op(std::string const &name, int i) const override;
  std::shared_ptr<SiteSet> clone() const override { return std::make_shared<Electron>(*this); }
};

class AutoMPO {
 public:
  explicit AutoMPO(SiteSet const &sites) : sites_(sites.clone()) {}
  AutoMPO &add(double coef, std::string const &op1, int i);
  AutoMPO &add(double coef, std::string const &op1, int i, std::string const &op2, int j);
  SiteSet const &sites() const { return *sites_; }

  // The site tensors before conversion: W[n] (n = 0..N-1) column-major over
  // (s, s', h_in, h_out) with link dims dims[n], dims[n+1] (links 0 and N
  // hold only the states done = 0 and start = 1), and each link's QN
  // sectors in state order (empty without QNs).
  struct Dense {
    std::vector<std::vector<double>> W;
    std::vector<int> dims;
    std::vector<QNBlocks> links;
  };
  Dense dense() const;

 private:
  struct Term {
    double coef;
    std::string a, b;  // b empty: one-site term
    int i, j;
  };
  std::shared_ptr<SiteSet> sites_;
  std::vector<Term> terms_;
};

// Bond (s1, s2) of a lattice graph with the coordinates of its ends
// (lattice/latticebond.h)
struct LatticeBond {
  int s1, s2;
  int x1, y1, x2, y2;
};
// Nx x Ny square lattice, sites n = (x-1) Ny + y, bonds in the order of
// the reference's squareLattice (lattice/square.h:24-48): per site the x
// bond, the y bond, and for YPeriodic (meaningful for Ny > 2) the wrap bond
std::vector<LatticeBond> squareLattice(int Nx, int Ny, bool yperiodic);

// W_1 (dag(s), s', l_1), W_n (dag(s), s', dag(l_{n-1}), l_n), W_N (dag(s),
// s', dag(l_{N-1})): the layout of the other MPOs of this layer
MPO toMPO(AutoMPO const &ampo);

}  // namespace itensor
