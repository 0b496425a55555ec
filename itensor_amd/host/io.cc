// Binary read / write of Index, ITensor and MPS in the stream format of the
// reference's write/read functions, so that files move between the two
// libraries:
//   Index    index.cc:711-745 : TagSet (MAX_TAGS = 4 tags of 7 chars, sorted
//            by their 8-byte integer value, then the prime level as int,
//            tagset.cc:265-285, smallstring.h:173-185), id (uint64), dim
//            (long), dir (Arrow, int), QN sectors as a vector of (QN, long)
//            (index.cc:561-586); a QN is 4 QNums of (name: 7 chars, val: int,
//            mod: int), active ones first in name order (qn.cc:88-105,
//            199-231, 499-508)
//   IndexSet indexset.cc:367-381 / tensor/range.h:57-75,284-298: count
//            (size_t) then (Index, stride: size_t) per index
//   ITensor  itensor.cc:768-806: IndexSet, scale (LogNum: double log|s| and
//            int sign, real.h:156-160; 1 here, USESCALE is off), StorageType
//            (int, itdata/task_types.h:454-472), then the storage: Dense =
//            vector<double> (itdata/dense.h:152-163), QDense = vector<BlOf>
//            (Block = InfArray<long>: count + longs; offset: long) then
//            vector<double> (itdata/qdense.h:236-248, util/readwrite.h:280-330)
//   MPS      mps/mps.cc:353-378: length (int), the N+2 site tensors of its
//            1-based store (the two ends empty), leftLim, rightLim (int)
// Integers are written in the host's byte order, as the reference does.
#include <algorithm>
#include <cstring>
#include <fstream>

#include "mps.h"

namespace itensor {

namespace {

constexpr int kMaxTags = 4;
constexpr size_t kTagChars = 7;
constexpr int kQNSize = 4;
enum StorageTypeId : int { ST_Null = 0, ST_DenseReal = 1, ST_QDenseReal = 6 };

template <typename T>
void put(std::ostream &s, T const &v) {
  s.write(reinterpret_cast<const char *>(&v), sizeof(T));
}
template <typename T>
T get(std::istream &s) {
  T v{};
  s.read(reinterpret_cast<char *>(&v), sizeof(T));
  if (!s) throw ITError("read: unexpected end of stream");
  return v;
}

// a tag or QN name as the reference's SmallString: 8 bytes, ordered by
// their value as a little-endian int64
struct Small {
  char c[8] = {};
  int64_t key() const {
    int64_t k;
    std::memcpy(&k, c, 8);
    return k;
  }
};
Small small(std::string const &t) {
  if (t.size() > kTagChars) throw ITError("write: tag or QN name longer than 7 characters: " + t);
  Small x;
  std::memcpy(x.c, t.data(), t.size());
  return x;
}
void putSmall(std::ostream &s, Small const &x) { s.write(x.c, kTagChars); }
Small getSmall(std::istream &s) {
  Small x;
  s.read(x.c, kTagChars);
  if (!s) throw ITError("read: unexpected end of stream");
  return x;
}

std::vector<std::string> splitTags(std::string const &t) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : t) {
    if (ch == ',') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else if (ch != ' ') {
      cur += ch;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

void writeQN(std::ostream &s, QN const &q) {
  std::vector<std::pair<Small, int>> nums;
  for (auto const &nv : q.store()) nums.push_back({small(nv.first), nv.second});
  if (int(nums.size()) > kQNSize) throw ITError("write: QN with more than 4 charges");
  std::sort(nums.begin(), nums.end(), [](auto const &a, auto const &b) { return a.first.key() < b.first.key(); });
  for (int i = 0; i < kQNSize; ++i) {
    if (i < int(nums.size())) {
      putSmall(s, nums[size_t(i)].first);
      put<int>(s, nums[size_t(i)].second);
      put<int>(s, 1);  // U(1): mod 1
    } else {
      putSmall(s, Small());
      put<int>(s, 0);
      put<int>(s, 0);  // inactive slot
    }
  }
}

QN readQN(std::istream &s) {
  QN q;
  for (int i = 0; i < kQNSize; ++i) {
    Small n = getSmall(s);
    const int v = get<int>(s), mod = get<int>(s);
    if (mod == 0) continue;  // inactive
    if (mod != 1) throw ITError("read: Z(n) quantum numbers (mod != 1) are not supported");
    q = q + QN(std::string(n.c, strnlen(n.c, kTagChars)), v);
  }
  return q;
}

}  // namespace

void write(std::ostream &s, Index const &I) {
  if (!I) throw ITError("Index::write: Index is default initialized");
  // TagSet: tags sorted by their integer value, empty slots last, then the
  // prime level
  std::vector<Small> tags;
  for (auto const &t : splitTags(I.tags())) tags.push_back(small(t));
  if (int(tags.size()) > kMaxTags) throw ITError("write: more than 4 tags on " + I.tags());
  std::sort(tags.begin(), tags.end(), [](Small const &a, Small const &b) { return a.key() < b.key(); });
  for (int i = 0; i < kMaxTags; ++i) putSmall(s, i < int(tags.size()) ? tags[size_t(i)] : Small());
  put<int>(s, I.primeLevel());
  put<uint64_t>(s, I.id());
  put<long>(s, I.dim());
  put<int>(s, int(I.dir()));
  put<size_t>(s, I.qnBlocks().size());
  for (auto const &qb : I.qnBlocks()) {
    writeQN(s, qb.first);
    put<long>(s, qb.second);
  }
}

Index readIndex(std::istream &s) {
  std::vector<std::string> tags;
  for (int i = 0; i < kMaxTags; ++i) {
    Small t = getSmall(s);
    if (t.c[0]) tags.emplace_back(t.c, strnlen(t.c, kTagChars));
  }
  const int plev = get<int>(s);
  const uint64_t id = get<uint64_t>(s);
  const long dim = get<long>(s);
  const int dir = get<int>(s);
  const size_t nq = get<size_t>(s);
  QNBlocks qns;
  for (size_t i = 0; i < nq; ++i) {
    QN q = readQN(s);
    qns.push_back({q, get<long>(s)});
  }
  std::string ts;
  for (auto const &t : tags) ts += (ts.empty() ? "" : ",") + t;
  return Index(id, dim, dir < 0 ? In : Out, plev < 0 ? 0 : plev, ts, qns);
}

void write(std::ostream &s, ITensor const &T) {
  IndexSet const &is = T.inds();
  put<size_t>(s, is.size());
  size_t stride = 1;
  for (auto const &I : is) {
    write(s, I);
    put<size_t>(s, stride);
    stride *= size_t(I.dim());
  }
  put<double>(s, 0.0);  // scale = LogNum(1): log|1|, sign +1
  put<int>(s, 1);
  if (!T) {
    put<int>(s, ST_Null);
    return;
  }
  if (T.isCombiner()) throw ITError("write: combiner storage is not written");
  std::vector<double> data = storageToHost(T);
  if (T.isQN()) {
    put<int>(s, ST_QDenseReal);
    auto const &offs = T.store().offsets;
    put<size_t>(s, offs.size());
    for (auto const &bo : offs) {
      put<size_t>(s, bo.block.size());
      for (int b : bo.block) put<long>(s, long(b));
      put<long>(s, long(bo.offset));
    }
  } else {
    put<int>(s, ST_DenseReal);
  }
  put<size_t>(s, data.size());
  s.write(reinterpret_cast<const char *>(data.data()), std::streamsize(sizeof(double) * data.size()));
}

ITensor readITensor(std::istream &s) {
  IndexSet is;
  const size_t n = get<size_t>(s);
  for (size_t i = 0; i < n; ++i) {
    is.push_back(readIndex(s));
    (void)get<size_t>(s);  // stride: column-major, recomputed
  }
  (void)get<double>(s);
  if (get<int>(s) != 1) throw ITError("read: ITensor scale other than +1 is not supported");
  const int type = get<int>(s);
  if (type == ST_Null) return ITensor(is, nullptr);
  BlockOffsets offs;
  if (type == ST_QDenseReal) {
    const size_t nb = get<size_t>(s);
    for (size_t b = 0; b < nb; ++b) {
      const size_t r = get<size_t>(s);
      Block blk(r);
      for (size_t q = 0; q < r; ++q) blk[q] = int(get<long>(s));
      offs.push_back({blk, int64_t(get<long>(s))});
    }
  } else if (type != ST_DenseReal) {
    throw ITError("read: storage type " + std::to_string(type) + " is not supported (real Dense and QDense are)");
  }
  const size_t nd = get<size_t>(s);
  std::vector<double> data(nd);
  s.read(reinterpret_cast<char *>(data.data()), std::streamsize(sizeof(double) * nd));
  if (!s) throw ITError("read: unexpected end of stream");
  if (type == ST_QDenseReal) return qdenseFromHost(is, offs, data);
  return denseFromHost(is, data);
}

void read(std::istream &s, ITensor &T) { T = readITensor(s); }

void writeToFile(std::string const &fname, ITensor const &T) {
  std::ofstream s(fname, std::ios::binary);
  if (!s.good()) throw ITError("Couldn't open file \"" + fname + "\" for writing");
  write(s, T);
}

ITensor readITensorFromFile(std::string const &fname) {
  std::ifstream s(fname, std::ios::binary);
  if (!s.good()) throw ITError("Couldn't open file \"" + fname + "\" for reading");
  return readITensor(s);
}

void write(std::ostream &s, MPS const &psi) {
  put<int>(s, psi.length());
  write(s, ITensor());  // A_[0]
  for (auto const &A : psi.tensors()) write(s, A);
  write(s, ITensor());  // A_[N+1]
  put<int>(s, psi.leftLim());
  put<int>(s, psi.rightLim());
}

MPS readMPS(std::istream &s) {
  const int N = get<int>(s);
  if (N < 0) throw ITError("read: negative MPS length");
  std::vector<ITensor> A;
  (void)readITensor(s);
  for (int j = 0; j < N; ++j) A.push_back(readITensor(s));
  (void)readITensor(s);
  const int ll = get<int>(s), rl = get<int>(s);
  return MPS(A, ll, rl);
}

void writeToFile(std::string const &fname, MPS const &psi) {
  std::ofstream s(fname, std::ios::binary);
  if (!s.good()) throw ITError("Couldn't open file \"" + fname + "\" for writing");
  write(s, psi);
}

MPS readMPSFromFile(std::string const &fname) {
  std::ifstream s(fname, std::ios::binary);
  if (!s.good()) throw ITError("Couldn't open file \"" + fname + "\" for reading");
  return readMPS(s);
}

}  // namespace itensor
