"""GPU: binary read / write of ITensor and MPS in the reference's stream
format.  The expected bytes are assembled here field by field from the
reference's write functions (index.cc:711-745 Index, tagset.cc:265-285
TagSet, smallstring.h:173-185 tag / QN-name characters, qn.cc:88-105,
499-508 QNum / QN, indexset.cc:367-381 and tensor/range.h:57-75,284-298
IndexSet, real.h:156-160 LogNum, itensor.cc:768-806 ITensor,
itdata/task_types.h:454-472 StorageType, itdata/dense.h:152-163 Dense,
itdata/qdense.h:236-248 QDense, util/readwrite.h:190-330 vectors, InfArray
and BlOf, mps/mps.cc:353-378 MPS); round trips are bit-exact."""
import os
import struct

import numpy as np
import pytest

from oracle import dmrg as OD
from tests.bridge import cten, oten

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H(torch_dev):
    from itensor_amd import host

    return host


def _small(s):
    b = s.encode().ljust(8, b"\0")
    return b, struct.unpack("<q", b)[0]


def _tagset(tags, plev):
    ts = sorted((_small(t) for t in tags), key=lambda x: x[1])
    out = b"".join(b[:7] for b, _ in ts) + b"\0" * 7 * (4 - len(ts))
    return out + struct.pack("<i", plev)


def _qn(q):
    nums = sorted(((_small(n), v) for n, v in q), key=lambda x: x[0][1])
    out = b""
    for i in range(4):
        if i < len(nums):
            (b, _), v = nums[i]
            out += b[:7] + struct.pack("<ii", v, 1)
        else:
            out += b"\0" * 7 + struct.pack("<ii", 0, 0)
    return out


def _index(I):
    out = _tagset([t for t in I.tags.split(",") if t], I.plev)
    out += struct.pack("<Qqi", I.id, I.dim, I.dir)
    qns = I.qns or []
    out += struct.pack("<Q", len(qns))
    for q, m in qns:
        out += _qn([(n, v) for n, v in q]) + struct.pack("<q", m)
    return out


def _itensor(T, data, offsets=None):
    out = struct.pack("<Q", len(T.inds))
    stride = 1
    for I in T.inds:
        out += _index(I) + struct.pack("<Q", stride)
        stride *= I.dim
    out += struct.pack("<di", 0.0, 1)
    if offsets is None:
        out += struct.pack("<i", 1)
    else:
        out += struct.pack("<i", 6) + struct.pack("<Q", len(offsets))
        for b, o in offsets:
            out += struct.pack("<Q", len(b)) + struct.pack("<%dq" % len(b), *b) + struct.pack("<q", o)
    return out + struct.pack("<Q", len(data)) + np.asarray(data, dtype="<f8").tobytes()


def test_dense_itensor_bytes(H):
    i = H.Index(0x1234567890ABCDEF, 3, 1, 0, "Link,l=3", [])
    j = H.Index(77, 2, -1, 2, "Site,S=1/2,n=4", [])
    data = np.arange(6, dtype=float) * 0.5 - 1.0
    T = H.dense([i, j], data)
    got = H.itensor_bytes(T)
    assert got == _itensor(T, data)
    back = H.itensor_from_bytes(got)
    assert [(x.id, x.dim, x.dir, x.plev, x.tags) for x in back.inds] == [(x.id, x.dim, x.dir, x.plev, x.tags)
                                                                        for x in T.inds]
    assert np.array_equal(back.store(), data)


def test_qdense_itensor_bytes_and_round_trip(H):
    rng = np.random.default_rng(3)
    a = H.Index(11, 5, 1, 0, "Link,l=1", [([("Nf", 1), ("Sz", -1)], 2), ([("Nf", 0)], 3)])
    b = H.Index(12, 4, -1, 1, "Site,n=2", [([("Nf", 1), ("Sz", -1)], 1), ([("Nf", 0)], 3)])
    T = H.random_itensor([a, b], [], 5)
    got = H.itensor_bytes(T)
    offs = [(list(x), o) for x, o in T.offsets]
    assert got == _itensor(T, T.store(), offs)
    back = H.itensor_from_bytes(got)
    assert [(tuple(x), o) for x, o in back.offsets] == [(tuple(x), o) for x, o in T.offsets]
    assert np.array_equal(back.store(), T.store())
    assert [x.qns for x in back.inds] == [x.qns for x in T.inds]


def test_mps_file_round_trip_after_dmrg(H, tmp_path):
    """A QN DMRG ground state written with MPS::write's layout and read
    back: every tensor, index and block list bit-exact, the limits kept; the
    product of the read tensors equals that of the written ones."""
    N = 8
    sites = OD.spin_half_sites(N, True)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)])
    E, recs, tens, secs = H.dmrg([cten(w) for w in Hm], [cten(a) for a in psi0.A],
                                 [{"MaxDim": 16, "Cutoff": 1e-12}] * 3, {})
    path = os.path.join(str(tmp_path), "psi.mps")
    H.write_mps(path, tens, 0, 2)
    raw = open(path, "rb").read()
    # length, the empty A_[0], the sites, the empty A_[N+1], the limits
    empty = struct.pack("<Q", 0) + struct.pack("<di", 0.0, 1) + struct.pack("<i", 0)
    want = struct.pack("<i", N) + empty
    for t in tens:
        want += _itensor(t, t.store(), [(list(x), o) for x, o in t.offsets])
    want += empty + struct.pack("<ii", 0, 2)
    assert raw == want
    back, ll, rl = H.read_mps(path)
    assert (ll, rl) == (0, 2)
    for a, b in zip(tens, back):
        assert [(x.id, x.plev, x.tags, x.dir, x.qns) for x in a.inds] == \
               [(x.id, x.plev, x.tags, x.dir, x.qns) for x in b.inds]
        assert [(tuple(x), o) for x, o in a.offsets] == [(tuple(x), o) for x, o in b.offsets]
        assert np.array_equal(a.store(), b.store())


def test_long_tag_is_an_error(H):
    i = H.Index(5, 2, 1, 0, "Link,averylongtag", [])
    with pytest.raises(Exception):
        H.itensor_bytes(H.dense([i], np.zeros(2)))
