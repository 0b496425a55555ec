// Args, QN, Index and block bookkeeping of the host layer (all integer /
// host-side work, as in the reference).
#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <sstream>

#include "itensor.h"

namespace itensor {

// ---------------------------------------------------------------- Args --
namespace {
template <typename T>
bool as(Args::Val const &v, T &out) {
  if (auto p = std::get_if<T>(&v)) {
    out = *p;
    return true;
  }
  return false;
}
}  // namespace

long Args::getInt(std::string const &k, long def) const {
  auto it = m_.find(k);
  if (it == m_.end()) return def;
  long l;
  double d;
  bool b;
  if (as(it->second, l)) return l;
  if (as(it->second, d)) return long(d);
  if (as(it->second, b)) return b ? 1 : 0;
  throw ITError("Args: key " + k + " is not a number");
}
long Args::getInt(std::string const &k) const {
  if (!defined(k)) throw ITError("Args: missing key " + k);
  return getInt(k, 0);
}
double Args::getReal(std::string const &k, double def) const {
  auto it = m_.find(k);
  if (it == m_.end()) return def;
  long l;
  double d;
  if (as(it->second, d)) return d;
  if (as(it->second, l)) return double(l);
  throw ITError("Args: key " + k + " is not a number");
}
double Args::getReal(std::string const &k) const {
  if (!defined(k)) throw ITError("Args: missing key " + k);
  return getReal(k, 0.0);
}
bool Args::getBool(std::string const &k, bool def) const {
  auto it = m_.find(k);
  if (it == m_.end()) return def;
  bool b;
  long l;
  if (as(it->second, b)) return b;
  if (as(it->second, l)) return l != 0;
  throw ITError("Args: key " + k + " is not a bool");
}
std::string Args::getString(std::string const &k, std::string const &def) const {
  auto it = m_.find(k);
  if (it == m_.end()) return def;
  std::string s;
  if (as(it->second, s)) return s;
  throw ITError("Args: key " + k + " is not a string");
}

// ------------------------------------------------------------------ QN --
void QN::set(std::string const &n, int v) {
  q_.clear();
  if (v != 0) q_.push_back({n, v});
}
void QN::add(std::string const &n, int v) {
  for (size_t k = 0; k < q_.size(); ++k)
    if (q_[k].first == n) {
      q_[k].second += v;
      if (q_[k].second == 0) q_.erase(q_.begin() + long(k));
      return;
    }
  if (v == 0) return;
  q_.push_back({n, v});
  std::sort(q_.begin(), q_.end());
}
int QN::val(std::string const &name) const {
  for (auto &p : q_)
    if (p.first == name) return p.second;
  return 0;
}
QN QN::operator+(QN const &o) const {
  QN r = *this;
  for (auto &p : o.q_) r.add(p.first, p.second);
  return r;
}
QN QN::operator-() const {
  QN r = *this;
  for (auto &p : r.q_) p.second = -p.second;
  return r;
}
std::string QN::str() const {
  std::ostringstream o;
  o << "QN(";
  for (size_t i = 0; i < q_.size(); ++i) o << (i ? "," : "") << q_[i].first << ":" << q_[i].second;
  o << ")";
  return o.str();
}

// --------------------------------------------------------------- Index --
namespace {
std::atomic<uint64_t> g_next_id{uint64_t(1) << 40};

std::string canon_tags(std::string const &t) {
  std::vector<std::string> v;
  std::string cur;
  for (char c : t) {
    if (c == ',') {
      if (!cur.empty()) v.push_back(cur);
      cur.clear();
    } else if (c != ' ') {
      cur += c;
    }
  }
  if (!cur.empty()) v.push_back(cur);
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  std::string r;
  for (size_t i = 0; i < v.size(); ++i) r += (i ? "," : "") + v[i];
  return r;
}
}  // namespace

uint64_t newIndexId() { return g_next_id.fetch_add(1); }

Index::Index(long dim, std::string const &tags) : id_(newIndexId()), dim_(dim), tags_(canon_tags(tags)) {}

Index::Index(QNBlocks qns, Arrow dir, std::string const &tags)
    : id_(newIndexId()), dir_(dir), tags_(canon_tags(tags)), qns_(std::move(qns)) {
  dim_ = 0;
  for (auto &q : qns_) dim_ += q.second;
}

Index::Index(uint64_t id, long dim, Arrow dir, int plev, std::string const &tags, QNBlocks qns)
    : id_(id), dim_(dim), dir_(dir), plev_(plev), tags_(canon_tags(tags)), qns_(std::move(qns)) {
  if (!qns_.empty()) {
    long d = 0;
    for (auto &q : qns_) d += q.second;
    if (d != dim_) throw ITError("Index: QN block sizes do not add up to dim");
  }
}

bool Index::hasTag(std::string const &t) const {
  std::string s = "," + tags_ + ",";
  return s.find("," + t + ",") != std::string::npos;
}

long Index::blockStart(long b) const {
  long s = 0;
  for (long i = 0; i < b; ++i) s += blocksize0(i);
  return s;
}

Index &Index::setTags(std::string const &t) {
  tags_ = canon_tags(t);
  return *this;
}

std::string Index::str() const {
  std::ostringstream o;
  o << "(" << dim_ << "|" << tags_ << "|" << id_ % 1000 << std::string(size_t(plev_), '\'') << ")";
  return o.str();
}

bool operator==(Index const &a, Index const &b) {
  return a.id() == b.id() && a.primeLevel() == b.primeLevel() && a.tags() == b.tags();
}

long dim(IndexSet const &is) {
  long d = 1;
  for (auto &i : is) d *= i.dim();
  return d;
}

int indexPosition(IndexSet const &is, Index const &I) {
  for (size_t k = 0; k < is.size(); ++k)
    if (is[k] == I) return int(k);
  return -1;
}

bool hasQNs(IndexSet const &is) {
  for (auto &i : is)
    if (i.hasQNs()) return true;
  return false;
}

bool blockLess(Block const &a, Block const &b) {
  return std::lexicographical_compare(a.rbegin(), a.rend(), b.rbegin(), b.rend());
}

QN calcDiv(IndexSet const &is, Block const &b) {
  QN q;
  for (size_t i = 0; i < is.size(); ++i) q = q + is[i].qn0(b[i]) * is[i].dir();
  return q;
}

// Every block of flux div in block order.  The charges are summed as small
// integer vectors (one entry per charge name): the first r-1 indices are
// walked as an odometer and the sectors of the last index with the missing
// charge are looked up, so the cost is the number of combinations of the
// first r-1 indices rather than of all r.
std::pair<BlockOffsets, int64_t> getBlockOffsets(IndexSet const &is, QN const &div) {
  BlockOffsets bofs;
  if (is.empty()) {
    bofs.push_back({Block(), 0});
    return {bofs, 1};
  }
  const int r = int(is.size());
  std::vector<std::string> names;
  auto addNames = [&](QN const &q) {
    for (auto const &p : q.store()) names.push_back(p.first);
  };
  addNames(div);
  for (auto const &I : is)
    for (long b = 0; b < I.nblock(); ++b)
      if (I.hasQNs()) addNames(I.qn0(b));
  std::sort(names.begin(), names.end());
  names.erase(std::unique(names.begin(), names.end()), names.end());
  using Charge = std::array<int, 4>;
  if (names.size() > 4) throw ITError("getBlockOffsets: more than four charge names");
  auto charge = [&](QN const &q, int sign) {
    Charge c{0, 0, 0, 0};
    for (auto const &p : q.store())
      c[size_t(std::lower_bound(names.begin(), names.end(), p.first) - names.begin())] = sign * p.second;
    return c;
  };
  std::vector<std::vector<Charge>> ch{size_t(r)};
  for (int j = 0; j < r; ++j) {
    Index const &I = is[size_t(j)];
    for (long b = 0; b < I.nblock(); ++b) ch[size_t(j)].push_back(I.hasQNs() ? charge(I.qn0(b), int(I.dir())) : Charge{});
  }
  std::map<Charge, std::vector<int>> last;
  for (int b = 0; b < int(ch[size_t(r - 1)].size()); ++b) last[ch[size_t(r - 1)][size_t(b)]].push_back(b);
  const Charge target = charge(div, 1);
  std::vector<Block> blocks;
  Block I(size_t(r), 0);
  while (true) {
    Charge need = target;
    for (int j = 0; j + 1 < r; ++j)
      for (int k = 0; k < 4; ++k) need[size_t(k)] -= ch[size_t(j)][size_t(I[size_t(j)])][size_t(k)];
    auto it = last.find(need);
    if (it != last.end())
      for (int b : it->second) {
        I[size_t(r - 1)] = b;
        blocks.push_back(I);
      }
    int j = 0;  // first index fastest
    for (; j + 1 < r; ++j) {
      if (++I[size_t(j)] < is[size_t(j)].nblock()) break;
      I[size_t(j)] = 0;
    }
    if (j + 1 >= r) break;
  }
  std::sort(blocks.begin(), blocks.end(), blockLess);
  int64_t total = 0;
  bofs.reserve(blocks.size());
  for (auto const &b : blocks) {
    int64_t d = 1;
    for (int j = 0; j < r; ++j) d *= is[size_t(j)].blocksize0(b[size_t(j)]);
    bofs.push_back({b, total});
    total += d;
  }
  return {bofs, total};
}

}  // namespace itensor
