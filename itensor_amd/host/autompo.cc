// Site sets and the finite-state-machine MPO builder (see autompo.h).
#include "autompo.h"

#include <algorithm>
#include <map>

namespace itensor {

namespace {

SiteOp zeros(int d) {
  SiteOp o;
  o.d = d;
  o.m.assign(size_t(d) * size_t(d), 0.0);
  return o;
}
void put(SiteOp &o, int s, int sp, double v) { o.m[size_t(s + o.d * sp)] = v; }  // <sp|O|s> = v
SiteOp diag(std::vector<double> const &v) {
  SiteOp o = zeros(int(v.size()));
  for (int s = 0; s < o.d; ++s) put(o, s, s, v[size_t(s)]);
  return o;
}
// (A B)(sp, s) = sum_t A(sp, t) B(t, s): B acts first
SiteOp product(SiteOp const &A, SiteOp const &B) {
  SiteOp C = zeros(A.d);
  for (int sp = 0; sp < A.d; ++sp)
    for (int s = 0; s < A.d; ++s) {
      double v = 0;
      for (int t = 0; t < A.d; ++t) v += A.m[size_t(t + A.d * sp)] * B.m[size_t(s + A.d * t)];
      C.m[size_t(s + A.d * sp)] = v;
    }
  C.flux = A.flux + B.flux;
  C.fermionic = A.fermionic != B.fermionic;
  return C;
}

std::string site_tag(std::string const &kind, int n) { return "Site," + kind + ",n=" + std::to_string(n); }

}  // namespace

// ---- SpinHalf (sites/spinhalf.h): Up = 0 (Sz +1), Dn = 1 (Sz -1) ----------
SpinHalf::SpinHalf(int N, bool conserveQNs) {
  for (int n = 1; n <= N; ++n) {
    if (conserveQNs)
      s_.emplace_back(QNBlocks{{QN("Sz", +1), 1}, {QN("Sz", -1), 1}}, Out, site_tag("S=1/2", n));
    else
      s_.emplace_back(2, site_tag("S=1/2", n));
  }
}

SiteOp SpinHalf::op(std::string const &name, int i) const {
  (void)i;
  if (name == "Id") return diag({1, 1});
  if (name == "Sz") return diag({0.5, -0.5});
  SiteOp o = zeros(2);
  if (name == "S+" || name == "Sp") {
    put(o, 1, 0, 1.0);
    o.flux = QN("Sz", +2);
    return o;
  }
  if (name == "S-" || name == "Sm") {
    put(o, 0, 1, 1.0);
    o.flux = QN("Sz", -2);
    return o;
  }
  throw ITError("SpinHalf: operator name " + name + " not recognized");
}

// ---- Electron (sites/electron.h): Emp 0, Up 1, Dn 2, UpDn 3 ----------------
Electron::Electron(int N, bool conserveQNs) {
  for (int n = 1; n <= N; ++n) {
    if (conserveQNs)
      s_.emplace_back(QNBlocks{{QN("Nf", 0, "Sz", 0), 1},
                               {QN("Nf", 1, "Sz", +1), 1},
                               {QN("Nf", 1, "Sz", -1), 1},
                               {QN("Nf", 2, "Sz", 0), 1}},
                      Out, site_tag("Elec", n));
    else
      s_.emplace_back(4, site_tag("Elec", n));
  }
}

SiteOp Electron::op(std::string const &name, int i) const {
  (void)i;
  constexpr int Emp = 0, Up = 1, Dn = 2, UpDn = 3;
  if (name == "Id") return diag({1, 1, 1, 1});
  if (name == "F") return diag({1, -1, -1, 1});
  if (name == "Fup") return diag({1, -1, 1, -1});
  if (name == "Fdn") return diag({1, 1, -1, -1});
  if (name == "Nup") return diag({0, 1, 0, 1});
  if (name == "Ndn") return diag({0, 0, 1, 1});
  if (name == "Nupdn") return diag({0, 0, 0, 1});
  if (name == "Ntot") return diag({0, 1, 1, 2});
  if (name == "Sz") return diag({0, 0.5, -0.5, 0});
  SiteOp o = zeros(4);
  // annihilation / creation; the down channel of C* carries (-1)^{n_up}
  const bool ferm = name[0] == 'C';
  const double sdn = ferm ? -1.0 : 1.0;
  if (name == "Cup" || name == "Aup") {
    put(o, Up, Emp, 1.0);
    put(o, UpDn, Dn, 1.0);
    o.flux = QN("Nf", -1, "Sz", -1);
  } else if (name == "Cdagup" || name == "Adagup") {
    put(o, Emp, Up, 1.0);
    put(o, Dn, UpDn, 1.0);
    o.flux = QN("Nf", +1, "Sz", +1);
  } else if (name == "Cdn" || name == "Adn") {
    put(o, Dn, Emp, 1.0);
    put(o, UpDn, Up, sdn);
    o.flux = QN("Nf", -1, "Sz", +1);
  } else if (name == "Cdagdn" || name == "Adagdn") {
    put(o, Emp, Dn, 1.0);
    put(o, Up, UpDn, sdn);
    o.flux = QN("Nf", +1, "Sz", -1);
  } else if (name == "S+" || name == "Sp") {
    put(o, Dn, Up, 1.0);
    o.flux = QN("Sz", +2);
  } else if (name == "S-" || name == "Sm") {
    put(o, Up, Dn, 1.0);
    o.flux = QN("Sz", -2);
  } else {
    throw ITError("Electron: operator name " + name + " not recognized");
  }
  o.fermionic = ferm;
  return o;
}

// ---- AutoMPO ----------------------------------------------------------------
AutoMPO &AutoMPO::add(double coef, std::string const &op1, int i) {
  if (i < 1 || i > sites_->length()) throw ITError("AutoMPO: site out of range");
  if (sites_->op(op1, i).fermionic) throw ITError("AutoMPO: a term with one fermionic operator is odd");
  terms_.push_back({coef, op1, "", i, i});
  return *this;
}

AutoMPO &AutoMPO::add(double coef, std::string const &op1, int i, std::string const &op2, int j) {
  const int N = sites_->length();
  if (i < 1 || i > N || j < 1 || j > N) throw ITError("AutoMPO: site out of range");
  const bool f1 = sites_->op(op1, i).fermionic, f2 = sites_->op(op2, j).fermionic;
  if (f1 != f2) throw ITError("AutoMPO: a term with one fermionic operator is odd");
  if (i == j) {
    terms_.push_back({coef, op1, op2, i, i});  // on-site product op1 * op2
  } else if (i < j) {
    terms_.push_back({coef, op1, op2, i, j});
  } else {  // order the sites; two fermionic operators anticommute
    terms_.push_back({f1 ? -coef : coef, op2, op1, j, i});
  }
  return *this;
}

AutoMPO::Dense AutoMPO::dense() const {
  SiteSet const &S = *sites_;
  const int N = S.length();
  if (N < 2) throw ITError("AutoMPO: needs at least two sites");
  const bool qn = S(1).hasQNs();
  constexpr int DONE = 0, START = 1;

  // open-term states, keyed by (first site, first operator)
  struct State {
    int i, jmax;
    std::string a;
    bool fermionic;
    QN q;
  };
  std::map<std::pair<int, std::string>, int> key;
  std::vector<State> st;
  for (auto const &t : terms_) {
    if (t.b.empty() || t.i == t.j) continue;
    auto k = std::make_pair(t.i, t.a);
    auto it = key.find(k);
    if (it == key.end()) {
      SiteOp A = S.op(t.a, t.i);
      key.emplace(k, int(st.size()));
      st.push_back({t.i, t.j, t.a, A.fermionic, -A.flux});
    } else {
      st[size_t(it->second)].jmax = std::max(st[size_t(it->second)].jmax, t.j);
    }
  }

  // per link: states in order (done, start, open states by first site),
  // then grouped into QN sectors in order of first appearance
  Dense D;
  std::vector<std::vector<int>> pos(size_t(N) + 1);  // pos[m][2 + state] = row on link m
  D.dims.resize(size_t(N) + 1);
  D.links.resize(size_t(N) + 1);
  for (int m = 0; m <= N; ++m) {
    std::vector<int> ids = {DONE, START};
    for (size_t s = 0; s < st.size(); ++s)
      if (st[s].i <= m && m < st[s].jmax) ids.push_back(2 + int(s));
    auto qof = [&](int id) { return id < 2 ? QN() : st[size_t(id - 2)].q; };
    std::vector<QN> sectors;
    std::vector<int> sec(ids.size());
    for (size_t x = 0; x < ids.size(); ++x) {
      const QN q = qof(ids[x]);
      auto f = std::find(sectors.begin(), sectors.end(), q);
      sec[x] = int(f - sectors.begin());
      if (f == sectors.end()) sectors.push_back(q);
    }
    std::vector<size_t> ord(ids.size());
    for (size_t x = 0; x < ord.size(); ++x) ord[x] = x;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return sec[a] < sec[b]; });
    pos[size_t(m)].assign(2 + st.size(), -1);
    for (size_t r = 0; r < ord.size(); ++r) pos[size_t(m)][size_t(ids[ord[r]])] = int(r);
    D.dims[size_t(m)] = int(ids.size());
    if (qn) {
      QNBlocks &lb = D.links[size_t(m)];
      for (size_t r = 0; r < ord.size(); ++r) {
        const QN &q = sectors[size_t(sec[ord[r]])];
        if (!lb.empty() && lb.back().first == q)
          ++lb.back().second;
        else
          lb.push_back({q, 1});
      }
    }
  }

  D.W.resize(size_t(N));
  for (int n = 1; n <= N; ++n) {
    const int d = int(S(n).dim()), Dl = D.dims[size_t(n - 1)], Dr = D.dims[size_t(n)];
    std::vector<double> &W = D.W[size_t(n - 1)];
    W.assign(size_t(d) * d * Dl * Dr, 0.0);
    auto add = [&](int hin, int hout, SiteOp const &O, double c) {
      for (int sp = 0; sp < d; ++sp)
        for (int s = 0; s < d; ++s)
          W[size_t(s + d * (sp + d * (hin + size_t(Dl) * hout)))] += c * O.m[size_t(s + d * sp)];
    };
    auto const &pl = pos[size_t(n - 1)], &pr = pos[size_t(n)];
    const SiteOp Id = S.op("Id", n);
    add(pl[DONE], pr[DONE], Id, 1.0);
    add(pl[START], pr[START], Id, 1.0);
    for (auto const &t : terms_) {
      if (t.b.empty() && t.i == n) add(pl[START], pr[DONE], S.op(t.a, n), t.coef);
      if (!t.b.empty() && t.i == n && t.j == n) add(pl[START], pr[DONE], product(S.op(t.a, n), S.op(t.b, n)), t.coef);
      if (!t.b.empty() && t.i < t.j && t.j == n) {
        const int s = key.at({t.i, t.a});
        add(pl[size_t(2 + s)], pr[DONE], S.op(t.b, n), t.coef);
      }
    }
    for (size_t s = 0; s < st.size(); ++s) {
      State const &x = st[s];
      if (x.i == n) {  // open: A (times F for a fermionic A, which acts first)
        const SiteOp A = S.op(x.a, n);
        add(pl[START], pr[2 + s], x.fermionic ? product(A, S.op("F", n)) : A, 1.0);
      } else if (x.i < n && n < x.jmax) {  // carry the string
        add(pl[2 + s], pr[2 + s], x.fermionic ? S.op("F", n) : Id, 1.0);
      }
    }
  }
  return D;
}

std::vector<LatticeBond> squareLattice(int Nx, int Ny, bool yperiodic) {
  yperiodic = yperiodic && Ny > 2;
  std::vector<LatticeBond> b;
  for (int n = 1; n <= Nx * Ny; ++n) {
    const int x = (n - 1) / Ny + 1, y = (n - 1) % Ny + 1;
    if (x < Nx) b.push_back({n, n + Ny, x, y, x + 1, y});
    if (Ny > 1) {
      if (y < Ny) b.push_back({n, n + 1, x, y, x, y + 1});
      if (yperiodic && y == 1) b.push_back({n, n + Ny - 1, x, y, x, y + Ny});
    }
  }
  return b;
}

MPO toMPO(AutoMPO const &ampo) {
  SiteSet const &S = ampo.sites();
  const AutoMPO::Dense D = ampo.dense();
  const int N = S.length();
  const bool qn = S(1).hasQNs();
  std::vector<Index> links(size_t(N) + 1);
  for (int m = 1; m < N; ++m) {
    const std::string tags = "Link,l=" + std::to_string(m);
    links[size_t(m)] = qn ? Index(D.links[size_t(m)], Out, tags) : Index(long(D.dims[size_t(m)]), tags);
  }
  auto make = [&](IndexSet const &is, std::vector<double> const &v) {
    return qn ? qdenseFromDense(is, v, QN()) : denseFromHost(is, v);
  };
  MPO H;
  for (int n = 1; n <= N; ++n) {
    Index s = S(n), sd = S(n), sp = S(n);
    sd.dag();
    sp.prime();
    const int d = int(s.dim()), Dl = D.dims[size_t(n - 1)], Dr = D.dims[size_t(n)];
    std::vector<double> const &W = D.W[size_t(n - 1)];
    if (n == 1) {  // row "start" (position 1 on the boundary link)
      std::vector<double> v(size_t(d) * d * Dr);
      for (int h = 0; h < Dr; ++h)
        for (int x = 0; x < d * d; ++x) v[size_t(x + d * d * h)] = W[size_t(x + d * d * (1 + size_t(Dl) * h))];
      H.push_back(make({sd, sp, links[1]}, v));
    } else if (n == N) {  // column "done" (position 0)
      Index l = links[size_t(N - 1)];
      l.dag();
      std::vector<double> v(size_t(d) * d * Dl);
      for (int h = 0; h < Dl; ++h)
        for (int x = 0; x < d * d; ++x) v[size_t(x + d * d * h)] = W[size_t(x + d * d * h)];
      H.push_back(make({sd, sp, l}, v));
    } else {
      Index l = links[size_t(n - 1)];
      l.dag();
      H.push_back(make({sd, sp, l, links[size_t(n)]}, W));
    }
  }
  return H;
}

}  // namespace itensor
