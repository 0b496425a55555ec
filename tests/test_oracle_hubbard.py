"""CPU tests: the fermion-model oracle against closed forms, and the host
layer's site sets and AutoMPO (dense site tensors, before any device
storage) against the oracle's exact Hamiltonians."""
import math

import numpy as np
import pytest

from oracle import hubbard as OH


# ---- the oracle itself ---------------------------------------------------------

def test_square_lattice_bond_counts():
    # Nbond = 2N - Ny (+0 periodic / -Nx open), square.h:30-31
    for Nx, Ny, yp in [(8, 4, True), (8, 4, False), (3, 2, True), (4, 1, False), (2, 3, True)]:
        N = Nx * Ny
        ypp = yp and Ny > 2
        assert len(OH.square_lattice(Nx, Ny, yp)) == 2 * N - Ny + (0 if ypp else -Nx)
    b = OH.square_lattice(2, 3, True)
    assert (1, 3) in b and (1, 4) in b and (1, 2) in b


@pytest.mark.parametrize("U", [0.0, 1.0, 4.0, 10.0])
def test_hubbard_dimer_closed_form(U):
    t = 1.0
    H = OH.hubbard_hamiltonian(2, [(1, 2)], t, U).toarray()
    assert np.allclose(H, H.T)
    e = OH.ground_energy(H, OH.electron_sector(2, 2, 0))
    assert abs(e - (U - math.sqrt(U * U + 16 * t * t)) / 2) < 1e-12


def test_hubbard_noninteracting_limit():
    # U = 0: ground energy = sum of the lowest single-particle levels per spin
    N, t = 5, 1.0
    bonds = [(1, 2), (2, 3), (3, 4), (4, 5), (1, 5), (2, 4)]
    Hsp = np.zeros((N, N))
    for i, j in bonds:
        Hsp[i - 1, j - 1] = Hsp[j - 1, i - 1] = -t
    eps = np.sort(np.linalg.eigvalsh(Hsp))
    H = OH.hubbard_hamiltonian(N, bonds, t, 0.0)
    for nup, ndn in [(2, 2), (3, 2), (1, 0), (3, 3)]:
        e = OH.ground_energy(H, OH.electron_sector(N, nup + ndn, nup - ndn))
        assert abs(e - (eps[:nup].sum() + eps[:ndn].sum())) < 1e-11


def test_hubbard_spin_flip_symmetry():
    N = 4
    bonds = OH.square_lattice(2, 2)
    H = OH.hubbard_hamiltonian(N, bonds, 1.0, 3.0)
    a = OH.ground_energy(H, OH.electron_sector(N, 3, 1))
    b = OH.ground_energy(H, OH.electron_sector(N, 3, -1))
    assert abs(a - b) < 1e-12


# ---- the host layer's MPO builder ----------------------------------------------

H_ = pytest.importorskip("itensor_amd.host", reason="host layer not built")


def mpo_matrix(W):
    """Full operator matrix (row = output) of AutoMPO.dense() site tensors
    W_n(s, s', h_in, h_out): boundary rows 'start' (1) and columns 'done' (0)."""
    T = W[0][:, :, 1, :]
    for w in W[1:]:
        d = w.shape[0]
        T = np.einsum("abh,cdhk->acbdk", T, w)
        n0 = T.shape[0] * d
        T = T.reshape((n0, n0, w.shape[3]), order="F")
    return T[:, :, 0].T


def check_flux(W, links, site_q):
    """Every nonzero element obeys q(h_out) = q(h_in) + q(s) - q(s')."""
    def expand(lb):
        out = []
        for q, n in lb:
            out += [dict(q)] * n
        return out

    def add(*qs):
        r = {}
        for sgn, q in qs:
            for k, v in q.items():
                r[k] = r.get(k, 0) + sgn * v
        return {k: v for k, v in r.items() if v}

    for n, w in enumerate(W):
        ql, qr = expand(links[n]), expand(links[n + 1])
        for s, sp, hi, ho in zip(*np.nonzero(w)):
            assert add((1, ql[hi]), (1, site_q[s]), (-1, site_q[sp])) == qr[ho], (n, s, sp, hi, ho)


ELEC_Q = [{}, {"Nf": 1, "Sz": 1}, {"Nf": 1, "Sz": -1}, {"Nf": 2}]


@pytest.mark.parametrize("bonds", [[(1, 2), (2, 3), (3, 4)], [(1, 2), (2, 3), (3, 4), (1, 4), (1, 3), (2, 4)]])
def test_autompo_hubbard_matches_ed(bonds):
    N, t, U = 4, 1.0, 4.0
    s = H_.electron_sites(N)
    a = H_.AutoMPO(s)
    for i, j in bonds:
        a.add(-t, "Cdagup", i, "Cup", j)
        a.add(-t, "Cdagup", j, "Cup", i)
        a.add(-t, "Cdagdn", i, "Cdn", j)
        a.add(-t, "Cdagdn", j, "Cdn", i)
    for i in range(1, N + 1):
        a.add(U, "Nupdn", i)
    W, dims, links = a.dense()
    M = mpo_matrix(W)
    ref = OH.hubbard_hamiltonian(N, bonds, t, U).toarray()
    assert np.max(np.abs(M - ref)) < 1e-13
    check_flux(W, links, ELEC_Q)


def test_autompo_mixed_terms_match_ed():
    # on-site products, reversed site order and density-density terms
    N = 4
    s = H_.electron_sites(N, conserve_qns=False)
    a = H_.AutoMPO(s)
    a.add(0.7, "Cdagdn", 4, "Cdn", 1)
    a.add(0.7, "Cdagdn", 1, "Cdn", 4)
    a.add(-0.3, "Ntot", 1, "Ntot", 3)
    a.add(1.5, "Cdagup", 2, "Cup", 2)  # = Nup on site 2
    a.add(0.2, "Sz", 2, "Sz", 3)
    W, _, _ = a.dense()
    M = mpo_matrix(W)
    c = OH._mode_ops(2 * N)
    n = [x.T @ x for x in c]
    ref = 0.7 * (c[7].T @ c[1] + c[1].T @ c[7])
    ref = ref - 0.3 * (n[0] + n[1]) @ (n[4] + n[5])
    ref = ref + 1.5 * n[2]
    ref = ref + 0.2 * (0.5 * (n[2] - n[3])) @ (0.5 * (n[4] - n[5]))
    ref = ref.toarray()
    assert np.max(np.abs(M - ref)) < 1e-13


def test_autompo_heisenberg_matches_ed():
    N = 5
    bonds = [(1, 2), (2, 3), (3, 4), (4, 5), (1, 5)]
    s = H_.spin_half_sites(N)
    a = H_.AutoMPO(s)
    for i, j in bonds:
        a.add(1.0, "Sz", i, "Sz", j)
        a.add(0.5, "S+", i, "S-", j)
        a.add(0.5, "S-", i, "S+", j)
    W, dims, links = a.dense()
    assert np.max(np.abs(mpo_matrix(W) - OH.heisenberg_hamiltonian(N, bonds))) < 1e-13
    check_flux(W, links, [{"Sz": 1}, {"Sz": -1}])


def test_autompo_rejects_odd_terms():
    a = H_.AutoMPO(H_.electron_sites(3))
    with pytest.raises(Exception):
        a.add(1.0, "Cup", 1)
    with pytest.raises(Exception):
        a.add(1.0, "Cup", 1, "Nup", 2)


@pytest.mark.parametrize("Nx,Ny,yp", [(8, 4, True), (3, 2, True), (2, 3, True), (4, 4, False), (5, 1, False)])
def test_square_lattice_matches_oracle(Nx, Ny, yp):
    assert [tuple(b) for b in H_.square_lattice(Nx, Ny, yp)] == OH.square_lattice(Nx, Ny, yp)


# ---- block enumeration of the host layer (no device) -----------------------------

def test_block_offsets_match_oracle():
    """getBlockOffsets of the host layer (charge lookup on the last index)
    against the oracle's exhaustive walk, on random index sets with one or
    two charges, mixed arrows and orders 1-4."""
    from oracle import itensor as OI
    rng = np.random.default_rng(7)
    names = [("Sz",), ("Nf", "Sz")]
    for trial in range(60):
        nm = names[trial % 2]
        r = 1 + trial % 4
        inds_o, inds_h = [], []
        for k in range(r):
            nsec = int(rng.integers(1, 6))
            secs = [({n: int(rng.integers(-2, 3)) for n in nm}, int(rng.integers(1, 4))) for _ in range(nsec)]
            d = int(rng.choice([-1, 1]))
            io = OI.Index([(OI.QN(q), m) for q, m in secs], f"i{k}", dir=d)
            inds_o.append(io)
            qns = [([(n, v) for n, v in sorted(q.items()) if v], m) for q, m in secs]
            inds_h.append(H_.Index(1000 + 10 * trial + k, sum(m for _, m in secs), d, 0, f"i{k}", qns))
        div = {n: int(rng.integers(-2, 3)) for n in nm}
        ref, rtot = OI.get_block_offsets(inds_o, OI.QN(div))
        got, gtot = H_.block_offsets(inds_h, [(n, v) for n, v in sorted(div.items()) if v])
        assert [tuple(b) for b, _ in got] == [tuple(b) for b, _ in ref], trial
        assert [o for _, o in got] == [o for _, o in ref] and gtot == rtot, trial
