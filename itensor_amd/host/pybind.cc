// Python bindings of the ITensor-compatible host layer (tests, benchmark).
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <sstream>

#include "autompo.h"
#include "mps.h"

namespace py = pybind11;
using namespace itensor;

namespace {

using QNList = std::vector<std::pair<std::string, int>>;

QN qnFrom(QNList const &l) {
  QN q;
  for (auto &p : l) q = q + QN(p.first, p.second);
  return q;
}

QNList qnTo(QN const &q) { return q.store(); }

Args argsFrom(py::dict const &d) {
  Args a;
  for (auto item : d) {
    std::string k = py::str(item.first);
    py::handle v = item.second;
    if (py::isinstance<py::bool_>(v))
      a.add(k, v.cast<bool>());
    else if (py::isinstance<py::int_>(v))
      a.add(k, v.cast<long>());
    else if (py::isinstance<py::float_>(v))
      a.add(k, v.cast<double>());
    else if (py::isinstance<py::str>(v))
      a.add(k, v.cast<std::string>());
    else
      throw ITError("unsupported Args value for key " + k);
  }
  return a;
}

Sweeps sweepsFrom(py::list const &l) {
  Sweeps s;
  for (auto item : l) {
    py::dict d = item.cast<py::dict>();
    SweepParams p;
    p.maxdim = d["MaxDim"].cast<long>();
    if (d.contains("MinDim")) p.mindim = d["MinDim"].cast<long>();
    if (d.contains("Cutoff")) p.cutoff = d["Cutoff"].cast<double>();
    if (d.contains("Niter")) p.niter = d["Niter"].cast<long>();
    if (d.contains("Noise")) p.noise = d["Noise"].cast<double>();
    s.push_back(p);
  }
  return s;
}

py::array_t<double> store_np(ITensor const &T) {
  auto h = storageToHost(T);
  py::array_t<double> a(h.size());
  std::copy(h.begin(), h.end(), a.mutable_data());
  return a;
}

py::dict recordDict(BondRecord const &r) {
  py::dict d;
  d["sweep"] = r.sweep;
  d["b"] = r.b;
  d["ha"] = r.ha;
  d["energy"] = r.energy;
  d["energy_fp64"] = r.energy_fp64;
  d["truncerr"] = r.truncerr;
  d["dim"] = r.dim;
  d["nprod"] = r.products;
  d["seconds"] = r.seconds;
  d["t_env"] = r.t_env;
  d["t_davidson"] = r.t_davidson;
  d["t_svd"] = r.t_svd;
  d["hpsi_flops"] = std::vector<double>(r.hpsi_flops.begin(), r.hpsi_flops.end());
  d["envs_device"] = r.envs_device;
  d["envs_host"] = r.envs_host;
  return d;
}

// LocalOp bound to tensors it co-owns (the reference's LocalOp holds
// non-owning pointers into LocalMPO, localop.h:46-49)
struct PyLocalOp {
  ITensor W1, W2, L, R, W12;
  bool fuse = true, fuseChecked = false;
  LocalOp op;
  PyLocalOp(ITensor w1, ITensor w2, ITensor l, ITensor r, bool f)
      : W1(std::move(w1)), W2(std::move(w2)), L(std::move(l)), R(std::move(r)), fuse(f) {
    op.update(&W1, &W2, &L, &R);
    if (fuse && op.isBulk()) W12 = W1 * W2;
  }
  void prepare(ITensor const &phi) {
    if (fuseChecked) return;
    if (fuse && op.isBulk()) op.setFused(&W12, phi.inds());
    fuseChecked = true;
  }
};

}  // namespace

PYBIND11_MODULE(_itensor_host, m) {
  m.doc() = "ITensor-compatible host layer of the MI355X DMRG H*psi engine";
  py::register_exception<ITError>(m, "ITError");
  // release the device context (plans, graphs, pools) at interpreter exit,
  // while the HIP runtime is still up; objects freed after that make no
  // HIP call
  m.def("shutdown", &itensor::shutdown);
  py::module_::import("atexit").attr("register")(py::cpp_function([]() { itensor::shutdown(); }));

  py::class_<Index>(m, "Index")
      .def(py::init([](uint64_t id, long dim, int dir, int plev, std::string const &tags,
                       std::vector<std::pair<QNList, long>> const &qns) {
             QNBlocks qb;
             for (auto &p : qns) qb.push_back({qnFrom(p.first), p.second});
             return Index(id, dim, Arrow(dir), plev, tags, qb);
           }),
           py::arg("id"), py::arg("dim"), py::arg("dir") = 1, py::arg("plev") = 0, py::arg("tags") = "",
           py::arg("qns") = std::vector<std::pair<QNList, long>>())
      .def_property_readonly("id", &Index::id)
      .def_property_readonly("dim", [](Index const &i) { return i.dim(); })
      .def_property_readonly("dir", [](Index const &i) { return int(i.dir()); })
      .def_property_readonly("plev", &Index::primeLevel)
      .def_property_readonly("tags", &Index::tags)
      .def_property_readonly("qns",
                             [](Index const &i) {
                               std::vector<std::pair<QNList, long>> r;
                               for (auto &p : i.qnBlocks()) r.push_back({qnTo(p.first), p.second});
                               return r;
                             })
      .def("prime", [](Index i, int inc) { return i.prime(inc); }, py::arg("inc") = 1)
      .def("dag", [](Index i) { return i.dag(); })
      .def("noprime", [](Index i) { return i.noPrime(); })
      .def("__eq__", [](Index const &a, Index const &b) { return a == b; })
      .def("__repr__", &Index::str);

  py::class_<ITensor>(m, "ITensor")
      .def(py::init<>())
      .def_property_readonly("inds", [](ITensor const &T) { return T.inds(); })
      .def_property_readonly("is_qn", &ITensor::isQN)
      .def_property_readonly("order", &ITensor::order)
      .def("nnz", &ITensor::nnz)
      .def("nnzblocks", &ITensor::nnzblocks)
      .def_property_readonly("offsets",
                             [](ITensor const &T) {
                               std::vector<std::pair<std::vector<int>, int64_t>> r;
                               for (auto &b : T.store().offsets) r.push_back({b.block, b.offset});
                               return r;
                             })
      .def("store", &store_np)
      .def("elt", &ITensor::elt)
      .def("__bool__", [](ITensor const &T) { return bool(T); })
      .def("__mul__", [](ITensor const &a, ITensor const &b) { return a * b; })
      .def("__mul__", [](ITensor const &a, double r) { return a * r; })
      .def("__rmul__", [](ITensor const &a, double r) { return a * r; })
      .def("__add__", [](ITensor const &a, ITensor const &b) { return a + b; })
      .def("__sub__", [](ITensor const &a, ITensor const &b) { return a - b; })
      .def("dag", [](ITensor const &a) { return dag(a); })
      .def("prime", [](ITensor const &a, int inc) { return prime(a, inc); }, py::arg("inc") = 1)
      .def("prime_index", [](ITensor const &a, Index const &I, int inc) { return prime(a, I, inc); },
           py::arg("I"), py::arg("inc") = 1)
      .def("noprime", [](ITensor const &a) { return noPrime(a); });

  m.def("dense", [](std::vector<Index> const &is, py::array_t<double, py::array::forcecast> a) {
    std::vector<double> v(a.data(), a.data() + a.size());
    return denseFromHost(is, v);
  }, "Dense ITensor from a flat column-major array");
  m.def("qdense", [](std::vector<Index> const &is, std::vector<std::pair<std::vector<int>, int64_t>> const &offs,
                     py::array_t<double, py::array::forcecast> a) {
    BlockOffsets bo;
    for (auto &p : offs) bo.push_back({p.first, p.second});
    std::vector<double> v(a.data(), a.data() + a.size());
    return qdenseFromHost(is, bo, v);
  }, "QDense ITensor from its block list and flat block store");
  m.def("norm", &norm);
  m.def("dot", &dotDag, "eltC(dag(A)*B)");
  m.def("remove_qns", &removeQNs);
  m.def("combiner", [](std::vector<Index> const &is, std::string const &tags) {
    Args a;
    a.add("Tags", tags);
    return combiner(is, a);
  }, py::arg("inds"), py::arg("tags") = "Link,CMB");
  m.def("diag_hermitian", [](ITensor const &rho, py::dict args) {
    ITensor U, D;
    auto spec = diag_hermitian(rho, U, D, argsFrom(args));
    return py::make_tuple(U, D, spec.eigs, spec.truncerr);
  });
  m.def("denmat_decomp", [](ITensor const &AA, ITensor A, ITensor B, bool fromleft, py::dict args) {
    auto spec = denmatDecomp(AA, A, B, fromleft ? Fromleft : Fromright, argsFrom(args));
    return py::make_tuple(A, B, spec.eigs, spec.truncerr);
  });
  m.def("svd", [](ITensor const &AA, std::vector<Index> const &uinds, py::dict args) {
    auto [U, D, V] = svd(AA, uinds, argsFrom(args));
    return py::make_tuple(U, D, V);
  }, py::arg("AA"), py::arg("uinds"), py::arg("args") = py::dict(), "svd(AA, {U indices}, args) -> (U, D, V)");
  m.def("svd_spec", [](ITensor const &AA, std::vector<Index> const &uinds, py::dict args) {
    ITensor U(uinds, nullptr), D, V;
    auto spec = svd(AA, U, D, V, argsFrom(args));
    return py::make_tuple(U, D, V, spec.eigs, spec.truncerr);
  });
  m.def("factor", [](ITensor const &T, std::vector<Index> const &ainds, py::dict args) {
    ITensor A(ainds, nullptr), B;
    auto spec = factor(T, A, B, argsFrom(args));
    return py::make_tuple(A, B, spec.eigs, spec.truncerr);
  }, py::arg("T"), py::arg("ainds"), py::arg("args") = py::dict());
  m.def("truncate", [](std::vector<double> P, long maxdim, long mindim, double cutoff, bool absc, bool rel,
                       bool respect) {
    Args a;
    a.add("RespectDegenerate", respect);
    auto r = truncate(P, maxdim, mindim, cutoff, absc, rel, a);
    return py::make_tuple(P, std::get<0>(r), std::get<1>(r), std::get<2>(r), std::get<3>(r));
  });
  m.def("davidson_bond", [](std::vector<ITensor> const &H, std::vector<ITensor> const &A, int b, py::dict args) {
    MPO mpo(H);
    MPS psi(A);
    LocalMPO PH(mpo);
    PH.position(b, psi);
    ITensor phi = psi(b) * psi(b + 1);
    DavidsonStats st;
    double e = davidson(PH, phi, argsFrom(args), &st);
    return py::make_tuple(e, phi, st.products);
  }, "davidson on the two-site problem of bond b (iterativesolvers_test.cc:43-65 pattern)");
  m.def("local_product", [](std::vector<ITensor> const &H, std::vector<ITensor> const &A, int b, ITensor const &phi,
                            int prec, bool fuse, int mode) {
    MPO mpo(H);
    MPS psi(A);
    LocalMPO PH(mpo);
    PH.position(b, psi);
    PH.setPrecision(prec);
    PH.setFusion(fuse);
    PH.setMode(HpsiMode(mode));
    PH.fuseOps(phi.inds());
    ITensor out;
    PH.product(phi, out);
    return out;
  }, py::arg("H"), py::arg("A"), py::arg("b"), py::arg("phi"), py::arg("prec") = 0, py::arg("fuse") = true,
     py::arg("mode") = int(HpsiMode::Graph));
  m.def("dmrg", [](std::vector<ITensor> const &H, std::vector<ITensor> const &A, py::list sweeps, py::dict args,
                   int left_lim, int right_lim) {
    MPS psi(A, left_lim, right_lim);
    std::vector<py::dict> recs;
    std::vector<double> secs;
    double e = dmrg(psi, H, sweepsFrom(sweeps), argsFrom(args),
                    [&](BondRecord const &r) { recs.push_back(recordDict(r)); }, &secs);
    return py::make_tuple(e, recs, psi.tensors(), secs);
  }, py::arg("H"), py::arg("A"), py::arg("sweeps"), py::arg("args"), py::arg("left_lim") = 0,
     py::arg("right_lim") = 2,
     "dmrg(H, psi, sweeps, args) (mps/dmrg.h:51-76); left_lim / right_lim are psi's orthogonality limits");
  m.def("write_itensor", [](std::string const &path, ITensor const &T) { writeToFile(path, T); },
        "writeToFile(path, T) in the reference's binary format (itensor.cc:768-806)");
  m.def("read_itensor", &readITensorFromFile, "readFromFile<ITensor>(path)");
  m.def("itensor_bytes", [](ITensor const &T) {
    std::ostringstream os(std::ios::binary);
    write(os, T);
    return py::bytes(os.str());
  }, "the bytes ITensor::write puts on a stream");
  m.def("itensor_from_bytes", [](py::bytes b) {
    std::istringstream is(std::string(b), std::ios::binary);
    return readITensor(is);
  });
  m.def("write_mps", [](std::string const &path, std::vector<ITensor> const &A, int ll, int rl) {
    writeToFile(path, MPS(A, ll, rl));
  }, py::arg("path"), py::arg("A"), py::arg("left_lim") = 0, py::arg("right_lim") = 2,
     "MPS::write to a file (mps/mps.cc:365-378)");
  m.def("read_mps", [](std::string const &path) {
    MPS psi = readMPSFromFile(path);
    return py::make_tuple(psi.tensors(), psi.leftLim(), psi.rightLim());
  }, "MPS::read from a file: (tensors, leftLim, rightLim)");
  m.def("position", [](std::vector<ITensor> const &A, int i, int left_lim, int right_lim, py::dict args) {
    MPS psi(A, left_lim, right_lim);
    position(psi, i, argsFrom(args));
    return py::make_tuple(psi.tensors(), psi.leftLim(), psi.rightLim());
  }, py::arg("A"), py::arg("i"), py::arg("left_lim"), py::arg("right_lim"), py::arg("args") = py::dict(),
     "MPS::position(i) (mps/mps.cc:667-735): (tensors, leftLim, rightLim)");
  m.def("svd_bond", [](std::vector<ITensor> const &A, int b, ITensor const &phi, bool fromleft, py::dict args) {
    MPS psi(A, fromleft ? b - 1 : b + 1, fromleft ? b + 1 : b + 2);
    auto spec = svdBond(psi, b, phi, fromleft ? Fromleft : Fromright, argsFrom(args));
    return py::make_tuple(psi(b), psi(b + 1), spec.eigs, spec.truncerr);
  }, py::arg("A"), py::arg("b"), py::arg("phi"), py::arg("fromleft"), py::arg("args") = py::dict(),
     "MPS::svdBond(b, phi, dir, args) (mps/mps_impl.h:21-99): (A_b, A_b+1, eigs, truncerr)");
  py::class_<PyLocalOp>(m, "LocalOp")
      .def(py::init<ITensor, ITensor, ITensor, ITensor, bool>(), py::arg("W1"), py::arg("W2"), py::arg("L"),
           py::arg("R"), py::arg("fuse") = true,
           "LocalOp(Op1, Op2, L, R) (localop.h:114-127); an empty ITensor() for L or R marks that end")
      .def("product", [](PyLocalOp &o, ITensor const &phi) {
        o.prepare(phi);
        ITensor out;
        o.op.product(phi, out);
        return out;
      })
      .def("flops", [](PyLocalOp &o, ITensor const &phi) {
        o.prepare(phi);
        return o.op.flops(phi);
      })
      .def("flops_by_form", [](PyLocalOp &o, ITensor const &phi) {
        o.prepare(phi);
        auto f = o.op.flopsByForm(phi);
        return std::vector<double>(f.begin(), f.end());
      }, "flops of one product by the form its kernels compute in: [f64, f32, bf16x3]")
      .def("delta_rho",
           [](PyLocalOp &o, ITensor const &AA, ITensor const &combine, bool fromleft) {
             auto r = o.op.deltaRho(AA, combine, fromleft ? Fromleft : Fromright);
             return py::make_tuple(r.first, r.second);
           },
           "deltaRho(AA, combiner, dir) (localop.h:375-406): (drho, trace)")
      .def("size", [](PyLocalOp &o) { return o.op.size(); })
      .def("set_precision", [](PyLocalOp &o, int p) { o.op.setPrecision(p); })
      .def("set_cover_output", [](PyLocalOp &o, bool on) { o.op.setCoverOutput(on); })
      .def("set_mode", [](PyLocalOp &o, int m) {
        if (m < 0 || m > 3) throw ITError("LocalOp.set_mode: mode must be 0..3");
        o.op.setMode(HpsiMode(m));
      })
      .def_property_readonly("fused", [](PyLocalOp const &o) { return bool(o.W12); });
  m.attr("HPSI_EAGER") = int(HpsiMode::Eager);
  m.attr("HPSI_PLANNED") = int(HpsiMode::Planned);
  m.attr("HPSI_GRAPH") = int(HpsiMode::Graph);
  m.attr("HPSI_PER_BLOCK") = int(HpsiMode::PerBlock);
  m.def("qn_task_list", [](std::vector<Index> const &Ais, std::vector<std::pair<std::vector<int>, int64_t>> const &Ao,
                           std::vector<Index> const &Bis, std::vector<std::pair<std::vector<int>, int64_t>> const &Bo) {
    BlockOffsets A, B;
    for (auto &x : Ao) A.push_back({x.first, x.second});
    for (auto &x : Bo) B.push_back({x.first, x.second});
    QNTaskList tl = qnTaskList(Ais, A, Bis, B);
    std::vector<std::pair<std::vector<int>, int64_t>> c;
    for (auto &b : tl.coffs) c.push_back({b.block, b.offset});
    py::dict d;
    d["inds"] = tl.N;
    d["coffs"] = c;
    d["csize"] = tl.csize;
    d["offs"] = tl.offs;
    d["cfirst"] = tl.cfirst;
    d["bda"] = tl.bda;
    d["bdb"] = tl.bdb;
    return d;
  }, "the block-pair task list of a QN contraction (structure only, host only)");
  m.def("block_offsets", [](std::vector<Index> const &is, QNList const &div) {
    auto [offs, total] = getBlockOffsets(is, qnFrom(div));
    std::vector<std::pair<std::vector<int>, int64_t>> r;
    for (auto &b : offs) r.push_back({b.block, b.offset});
    return std::make_pair(r, total);
  }, py::arg("inds"), py::arg("div") = QNList(), "getBlockOffsets: the blocks of flux div in block order, host only");
  m.def("random_itensor", [](std::vector<Index> const &is, QNList const &div, uint64_t seed) {
    return randomITensor(qnFrom(div), is, seed);
  }, py::arg("inds"), py::arg("div") = QNList(), py::arg("seed") = 1);
  m.def("from_dense", [](std::vector<Index> const &is, py::array_t<double, py::array::forcecast | py::array::f_style> a,
                         QNList const &div) {
    std::vector<double> v(a.data(), a.data() + a.size());
    if (!hasQNs(is)) return denseFromHost(is, v);
    return qdenseFromDense(is, v, qnFrom(div));
  }, py::arg("inds"), py::arg("array"), py::arg("div") = QNList(),
     "ITensor from a dense array (column-major); QDense of flux div when the indices carry QNs");
  py::class_<SiteSet, std::shared_ptr<SiteSet>>(m, "SiteSet")
      .def("__call__", [](SiteSet const &s, int i) { return s(i); })
      .def("__len__", &SiteSet::length)
      .def("op_matrix",
           [](SiteSet const &s, std::string const &name, int i) {
             SiteOp o = s.op(name, i);  // returned as M[sp, s] = <sp|O|s>
             py::array_t<double> a({o.d, o.d});
             auto r = a.mutable_unchecked<2>();
             for (int sp = 0; sp < o.d; ++sp)
               for (int x = 0; x < o.d; ++x) r(sp, x) = o.m[size_t(x + o.d * sp)];
             return py::make_tuple(a, qnTo(o.flux), o.fermionic);
           });
  m.def("square_lattice", [](int Nx, int Ny, bool yperiodic) {
    std::vector<std::pair<int, int>> r;
    for (auto &b : squareLattice(Nx, Ny, yperiodic)) r.push_back({b.s1, b.s2});
    return r;
  }, py::arg("Nx"), py::arg("Ny"), py::arg("yperiodic") = false);
  m.def("spin_half_sites", [](int N, bool qn) { return std::shared_ptr<SiteSet>(std::make_shared<SpinHalf>(N, qn)); },
        py::arg("N"), py::arg("conserve_qns") = true);
  m.def("electron_sites", [](int N, bool qn) { return std::shared_ptr<SiteSet>(std::make_shared<Electron>(N, qn)); },
        py::arg("N"), py::arg("conserve_qns") = true);
  py::class_<AutoMPO>(m, "AutoMPO")
      .def(py::init([](std::shared_ptr<SiteSet> const &s) { return AutoMPO(*s); }))
      .def("add", [](AutoMPO &a, double c, std::string const &o, int i) { a.add(c, o, i); })
      .def("add", [](AutoMPO &a, double c, std::string const &o1, int i, std::string const &o2, int j) {
        a.add(c, o1, i, o2, j);
      })
      .def("dense",
           [](AutoMPO const &a) {  // host-only: W_n(s, s', h_in, h_out) arrays and link sectors
             AutoMPO::Dense D = a.dense();
             py::list W, links;
             for (size_t n = 0; n < D.W.size(); ++n) {
               const py::ssize_t d = a.sites()(int(n) + 1).dim();
               py::array_t<double, py::array::f_style> w({d, d, py::ssize_t(D.dims[n]), py::ssize_t(D.dims[n + 1])});
               std::copy(D.W[n].begin(), D.W[n].end(), w.mutable_data());
               W.append(w);
             }
             for (auto &lb : D.links) {
               py::list l;
               for (auto &q : lb) l.append(py::make_tuple(qnTo(q.first), q.second));
               links.append(l);
             }
             return py::make_tuple(W, D.dims, links);
           })
      .def("to_mpo", [](AutoMPO const &a) { return toMPO(a); });

  m.def("set_stream", [](uintptr_t s) { setStream(reinterpret_cast<void *>(s)); });
  m.def("synchronize", []() { check(itx_synchronize(ctx()), "itx_synchronize"); });
  m.def("set_contract_precision", &setContractPrecision);
  m.def("set_autotune", [](bool on) { check(itx_set_autotune(ctx(), on ? 1 : 0), "itx_set_autotune"); });
}
