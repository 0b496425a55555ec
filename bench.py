"""Benchmark of the DMRG two-site H*psi hot path on one MI355X.

Workload (one "step"): one dense two-site effective-Hamiltonian product
H*psi = noPrime(((psi*L)*W1)*W2*R) with psi in R^{chi x 2 x 2 x chi},
L, R in R^{chi x 5 x chi}, W1 = W2 = the bulk Heisenberg MPO tensor
(5 x 2 x 2 x 5), chi = 1024, entries of psi, L, R i.i.d. N(0,1) from a fixed
seed (synthetic data), all operands resident in HBM before timing.
The contractions run through the libitx C ABI (plans built once), in the
order LocalOp::product runs them: psi*L, then the MPO pair pre-contracted
(W12 = W1*W2, same flop count and result layout), then *R (--chain 4stage
runs the reference's four separate stages instead).

value = algorithmic FP64-equivalent flops (80 chi^3 + 800 chi^2 per H*psi)
        / wall time, whole job (one GPU).

Further legs on the same line: "qn_hpsi" -- the block-sparse H*psi at
chi = 1024 over 31 Sz sectors in every execution mode (grouped chain as a
HIP graph / launched directly / eager, one launch per block pair, FP32 and
BF16x3 forms); "dmrg" -- QN Heisenberg DMRG N=100 with the maxdim ramp to
2048 (cutoff 1e-10) in the FP64, FP32 and BF16x3 H*psi forms: seconds per
sweep, energies and their differences.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--chi 1024]
                       [--no-cpu] [--no-qn] [--dmrg N:maxdims|''] [--dmrg-forms f64,f32,bf16x3]
                       [--prec f64|f32|bf16x3] [--chain fused|4stage]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DMRG sec/sweep + Hψ TFLOP/s, mixed vs FP64 on 1 MI355X at |ΔE|≤1e-9"
# TFLOP/s dense MFMA (datasheet); the BF16x3 form runs 3 BF16 products per
# FP64-equivalent product, so its FP64-equivalent ceiling is 2500/3
PEAK = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0 / 3}
SEED = 20261202 + 2


def stage_dims(chi: int, d: int = 2, k: int = 5):
    psi = (chi, d, d, chi)
    Ld = (chi, k, chi)
    W = (k, d, d, k)
    T1 = (d, d, chi, k, chi)
    T2 = (d, chi, chi, d, k)
    T3 = (chi, chi, d, d, k)
    return [(psi, Ld, T1), (T1, W, T2), (T2, W, T3), (T3, Ld, psi)]


# labels of the four stages as ITensor's computeLabels assigns them
STAGES = [
    ((-1, 2, 3, 4), (-1, 5, 6), (2, 3, 4, 5, 6)),
    ((-1, 3, 4, -2, 5), (-2, -1, 6, 7), (3, 4, 5, 6, 7)),
    ((-1, 3, 4, 5, -2), (-2, -1, 6, 7), (3, 4, 5, 6, 7)),
    ((-1, 3, 4, 5, -2), (-1, -2, 6), (3, 4, 5, 6)),
]
# fused chain of LocalOp::product: S1, then T1 x W12 with W12 = W1*W2
# (h0,s1,s1',s2,s2',h2), then S4 -- same result layout and flop count
FUSED = [
    STAGES[0],
    ((-1, -2, 3, -3, 4), (-3, -1, 5, -2, 6, 7), (3, 4, 5, 6, 7)),
    STAGES[3],
]


def fused_dims(chi: int, d: int = 2, k: int = 5):
    psi = (chi, d, d, chi)
    Ld = (chi, k, chi)
    T1 = (d, d, chi, k, chi)
    W12 = (k, d, d, d, d, k)
    T3 = (chi, chi, d, d, k)
    return [(psi, Ld, T1), (T1, W12, T3), (T3, Ld, psi)]


def heisenberg_w() -> np.ndarray:
    """Bulk Heisenberg MPO tensor W(h_in, s, s', h_out) (J = Jz = 1)."""
    Id = np.eye(2)
    Sz = np.diag([0.5, -0.5])
    Sp = np.zeros((2, 2)); Sp[1, 0] = 1.0
    Sm = np.zeros((2, 2)); Sm[0, 1] = 1.0
    W = np.zeros((5, 2, 2, 5))
    W[0, :, :, 0] = Id; W[1, :, :, 1] = Id
    W[2, :, :, 0] = Sz; W[1, :, :, 2] = Sz
    W[3, :, :, 0] = Sm; W[1, :, :, 3] = 0.5 * Sp
    W[4, :, :, 0] = Sp; W[1, :, :, 4] = 0.5 * Sm
    return W


def hpsi_flops(chi: int, d: int = 2, k: int = 5) -> float:
    return 4.0 * d * d * k * chi ** 3 + 4.0 * d ** 3 * k * k * chi ** 2


def cpu_baseline(psi, L, W, R, budget_s: float = 20.0):
    """The oracle's H*psi (numpy tensordot over the host BLAS) on the host
    cores: 1 warm-up, then repeats until ~budget_s, median reported."""
    from oracle.dense import hpsi_dense

    try:
        from threadpoolctl import threadpool_info

        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    ref = hpsi_dense(psi, L, W, W, R)
    first = time.perf_counter() - t0
    times = []
    reps = max(1, min(5, int(budget_s / max(first, 1e-3))))
    for _ in range(reps):
        t0 = time.perf_counter()
        hpsi_dense(psi, L, W, W, R)
        times.append(time.perf_counter() - t0)
    return ref, float(np.median(times)), threads, reps


# bond sectors q = 2 Sz in {-30..30} of the chi = 1024 Sz-conserving
# Heisenberg bond at the chain midpoint (binomial weights, max 120, min 1)
SECTORS_1024 = [1, 1, 1, 1, 1, 1, 3, 7, 14, 25, 41, 60, 81, 101, 114, 120, 114, 101, 81, 60, 41, 25, 14, 7, 3,
                1, 1, 1, 1, 1, 1]


def qn_operands(H, sectors, seed):
    """Block-sparse two-site operands: phi(dag(l), s1, s2, r), L(l, hl,
    dag(l)'), R(dag(r), dag(hr), r'), W(dag(s), s', dag(h_in), h_out);
    site {+1: 1, -1: 1}, MPO link {0: 3, -2: 1, +2: 1}; flux-0 blocks of
    phi, L, R filled with N(0,1)."""
    def sz(v):
        return [("Sz", int(v))]

    ids = iter(range(7001, 7100))

    def idx(qns, tags):
        return H.Index(next(ids), sum(d for _, d in qns), 1, 0, tags, qns)

    bond = [(sz(q), d) for q, d in zip(range(-30, 31, 2), sectors)]
    l, r = idx(bond, "Link,l=49"), idx(bond, "Link,l=51")
    site = [(sz(1), 1), (sz(-1), 1)]
    s1, s2 = idx(site, "Site,n=50"), idx(site, "Site,n=51")
    hq = [(sz(0), 3), (sz(-2), 1), (sz(2), 1)]
    hl, hm, hr = idx(hq, "Link,h=49"), idx(hq, "Link,h=50"), idx(hq, "Link,h=51")
    W = heisenberg_w().transpose(1, 2, 0, 3)  # (s, s', h_in, h_out)
    W1 = H.from_dense([s1.dag(), s1.prime(), hl.dag(), hm], W)
    W2 = H.from_dense([s2.dag(), s2.prime(), hm.dag(), hr], W)
    phi = H.random_itensor([l.dag(), s1, s2, r], [], seed)
    L = H.random_itensor([l, hl, l.dag().prime()], [], seed + 1)
    R = H.random_itensor([r.dag(), hr.dag(), r.prime()], [], seed + 2)
    return phi, L, W1, W2, R


def qn_hpsi(steps: int, warmup: int):
    """Block-sparse H*psi (chi = 1024 over 31 Sz sectors) through
    LocalOp::product in each execution mode: the grouped-GEMM chain
    replayed as a HIP graph, the same chain launched directly, eager
    operator* products, and one launch per block pair (the reference's
    loop).  Wall time per product over `steps` products after `warmup`."""
    import torch

    from itensor_amd import host as H

    phi, L, W1, W2, R = qn_operands(H, SECTORS_1024, 20261202 + 3)
    modes = [("graph", H.HPSI_GRAPH, 0), ("planned", H.HPSI_PLANNED, 0), ("eager", H.HPSI_EAGER, 0),
             ("per_block", H.HPSI_PER_BLOCK, 0), ("graph_f32", H.HPSI_GRAPH, 1),
             ("graph_bf16x3", H.HPSI_GRAPH, 2)]
    res, ref = {}, None
    for name, mode, prec in modes:
        op = H.LocalOp(W1, W2, L, R, True)
        op.set_mode(mode)
        op.set_precision(prec)
        flops = op.flops(phi) if mode != H.HPSI_EAGER else res["graph"]["flops"]
        t_plan = time.perf_counter()
        out = op.product(phi)
        torch.cuda.synchronize()
        t_plan = time.perf_counter() - t_plan
        for _ in range(warmup):
            op.product(phi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            op.product(phi)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        st = out.store()
        r = {"ms_per_hpsi": round(ms, 4), "tflops": round(flops / (ms * 1e-3) / 1e12, 3), "flops": flops,
             "first_call_ms": round(t_plan * 1e3, 2)}
        if ref is None:
            ref = (st, out.offsets)
            r["blocks_out"] = len(out.offsets)
        else:
            r["same_block_list"] = out.offsets == ref[1]
            r["rel_err_vs_graph_f64"] = float(np.max(np.abs(st - ref[0])) / np.max(np.abs(ref[0])))
        res[name] = r
    res["blocks_phi"] = len(phi.offsets)
    res["speedup_grouped_vs_per_block"] = round(res["per_block"]["ms_per_hpsi"] / res["graph"]["ms_per_hpsi"], 2)
    return res


def heisenberg_qn_chain(H, N: int):
    """Sz-conserving spin-1/2 Heisenberg chain (J = Jz = 1, open ends):
    site indices, the MPO (bulk W(dag(s), s', dag(h_in), h_out), first and
    last site contracted with the start / end states of the MPO link,
    sample/Heisenberg.h:76-116) and the Neel product state with link QNs
    that give every site tensor flux 0."""
    def sz(v):
        return [("Sz", int(v))]

    ids = iter(range(100001, 100001 + 4 * N + 8))

    def idx(qns, tags):
        return H.Index(next(ids), sum(d for _, d in qns), 1, 0, tags, qns)

    sites = [idx([(sz(1), 1), (sz(-1), 1)], f"Site,S=1/2,n={n}") for n in range(1, N + 1)]
    hq = [(sz(0), 3), (sz(-2), 1), (sz(2), 1)]
    hl = [idx(hq, f"Link,H,l={n}") for n in range(N + 1)]
    Wb = heisenberg_w().transpose(1, 2, 0, 3)  # (s, s', h_in, h_out)
    mpo = []
    for n in range(1, N + 1):
        s, row, col = sites[n - 1], hl[n - 1].dag(), hl[n]
        if n == 1:
            W, inds = Wb[:, :, 1, :], [s.dag(), s.prime(), col]
        elif n == N:
            W, inds = Wb[:, :, :, 0], [s.dag(), s.prime(), row]
        else:
            W, inds = Wb, [s.dag(), s.prime(), row, col]
        mpo.append(H.from_dense(inds, np.ascontiguousarray(W)))
    # Neel state: link l_n carries -(sum of the site charges 1..n)
    A, q, prev = [], 0, None
    for n in range(1, N + 1):
        st = (n - 1) % 2  # 0 = Up (Sz +1), 1 = Dn
        q -= 1 if st == 0 else -1
        s = sites[n - 1]
        v = np.zeros(2)
        v[st] = 1.0
        if n < N:
            link = idx([(sz(q), 1)], f"Link,l={n}")
        if n == 1:
            A.append(H.from_dense([s, link], v.reshape(2, 1)))
        elif n == N:
            A.append(H.from_dense([prev.dag(), s], v.reshape(1, 2)))
        else:
            A.append(H.from_dense([prev.dag(), s, link], v.reshape(1, 2, 1)))
        prev = link if n < N else None
    return mpo, A


def hubbard_cylinder(H, Nx: int, Ny: int, t: float = 1.0, U: float = 4.0, yperiodic: bool = True):
    """Hubbard model on the Nx x Ny square lattice (YPeriodic cylinder, site
    order of squareLattice) with (Nf, Sz) conserved: the MPO from the host
    layer's AutoMPO (-t c+_i,s c_j,s + h.c. on every bond, U n_up n_dn on
    every site) and a half-filled product state (Up, Dn, Up, ...) whose link
    QNs give every site tensor but the last flux 0; the last one carries the
    state's total charge (Nf = N, Sz = 0)."""
    N = Nx * Ny
    sites = H.electron_sites(N)
    ampo = H.AutoMPO(sites)
    for (i, j) in H.square_lattice(Nx, Ny, yperiodic):
        for c, a in (("Cdagup", "Cup"), ("Cdagdn", "Cdn")):
            ampo.add(-t, c, i, a, j)
            ampo.add(-t, c, j, a, i)
    for i in range(1, N + 1):
        ampo.add(U, "Nupdn", i)
    mpo = ampo.to_mpo()
    ids = iter(range(300001, 300001 + N + 8))
    A, nf, sz, prev = [], 0, 0, None
    for n in range(1, N + 1):
        st = 1 if (n - 1) % 2 == 0 else 2  # Up, Dn
        nf += 1
        sz += 1 if st == 1 else -1
        s = sites(n)
        v = np.zeros(4)
        v[st] = 1.0
        if n < N:
            q = [(k, x) for k, x in (("Nf", -nf), ("Sz", -sz)) if x]
            link = H.Index(next(ids), 1, 1, 0, f"Link,l={n}", [(q, 1)])
        if n == 1:
            A.append(H.from_dense([s, link], v.reshape(4, 1)))
        elif n == N:
            tot = [(k, x) for k, x in (("Nf", nf), ("Sz", sz)) if x]
            A.append(H.from_dense([prev.dag(), s], v.reshape(1, 4), tot))
        else:
            A.append(H.from_dense([prev.dag(), s, link], v.reshape(1, 4, 1)))
        prev = link if n < N else None
    return mpo, A


DMRG_MAXDIM = [10, 20, 50, 100, 200, 400, 800, 1200, 1600, 2048]


def dmrg_leg(N: int, maxdims, prec: int, mode: int, final_fp64: int = 0):
    """QN Heisenberg DMRG (N sites, Neel start, one sweep per maxdim entry,
    cutoff 1e-10, Davidson niter 2) in one H*psi form: wall time of each
    sweep (everything included) and how it splits over the phases of the
    bond steps (environment update, davidson, svdBond)."""
    import torch

    from itensor_amd import host as H

    mpo, A = heisenberg_qn_chain(H, N)
    sweeps = [{"MaxDim": int(m), "Cutoff": 1e-10, "Niter": 2} for m in maxdims]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E, recs, _, secs = H.dmrg(mpo, A, sweeps, {"HpsiPrecision": prec, "HpsiMode": mode, "SyncTiming": True,
                                               "FinalFP64Sweeps": final_fp64})
    wall = time.perf_counter() - t0
    last = [r for r in recs if r["sweep"] == len(maxdims)]
    return {"energy": E, "wall_s": round(wall, 3), "sec_per_sweep": [round(x, 4) for x in secs],
            "max_dim": [max(r["dim"] for r in recs if r["sweep"] == k + 1) for k in range(len(maxdims))],
            "truncerr_max_last_sweep": max(r["truncerr"] for r in last),
            "last_sweep_split_s": {"env": round(sum(r["t_env"] for r in last), 4),
                                   "davidson": round(sum(r["t_davidson"] for r in last), 4),
                                   "svd": round(sum(r["t_svd"] for r in last), 4)},
            "hpsi_products_last_sweep": sum(r["nprod"] for r in last)}


HUBBARD_MAXDIM = [50, 100, 200, 400, 400, 400, 400, 400]  # ramp, then converged at 400


def hubbard_leg(Nx: int, Ny: int, maxdims, prec: int, final_fp64: int = 0):
    """(Nf, Sz)-conserving Hubbard DMRG on the Nx x Ny YPeriodic cylinder
    (t = 1, U = 4, half filling, cutoff 1e-9, niter 2, noise 1e-6 / 1e-7 on
    the first two sweeps) in one H*psi form: wall time per sweep and the
    phase split of the last sweep."""
    import torch

    from itensor_amd import host as H

    mpo, A = hubbard_cylinder(H, Nx, Ny, 1.0, 4.0, True)
    noise = [1e-6, 1e-7] + [0.0] * max(0, len(maxdims) - 2)
    sweeps = [{"MaxDim": int(m), "Cutoff": 1e-9, "Niter": 2, "Noise": z} for m, z in zip(maxdims, noise)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E, recs, _, secs = H.dmrg(mpo, A, sweeps, {"HpsiPrecision": prec, "SyncTiming": True,
                                               "FinalFP64Sweeps": final_fp64})
    wall = time.perf_counter() - t0
    last = [r for r in recs if r["sweep"] == len(maxdims)]
    return {"energy": E, "wall_s": round(wall, 3), "sec_per_sweep": [round(x, 4) for x in secs],
            "max_dim": [max(r["dim"] for r in recs if r["sweep"] == k + 1) for k in range(len(maxdims))],
            "truncerr_max_last_sweep": max(r["truncerr"] for r in last),
            "last_sweep_split_s": {"env": round(sum(r["t_env"] for r in last), 4),
                                   "davidson": round(sum(r["t_davidson"] for r in last), 4),
                                   "svd": round(sum(r["t_svd"] for r in last), 4)}}


def measured_traffic(chi: int, prec: str, stage: str):
    """HBM bytes per launch of a stage as rocprofv3 FETCH_SIZE + WRITE_SIZE
    measured them (copy-calibrated) -- read from the committed profile
    summary of the chi = 1024 FP64 chain; None for other workloads."""
    if chi != 1024 or prec != "f64":
        return None
    path = os.path.join(ROOT, "profiles", "pmc_dense_hpsi_chi1024.md")
    try:
        import re

        m = re.search(r"traffic_bytes_per_launch %s = (\d+)" % stage, open(path).read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def copy_calibration(gib: float):
    """A plain device-to-device copy of known size (reads and writes
    `gib` GiB per launch), the reference point for HBM byte counters."""
    import torch

    import itensor_amd  # noqa: F401

    n = int(gib * (1 << 30) / 8)
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": n * 8, "bytes_written": n * 8, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chi", type=int, default=1024)
    ap.add_argument("--prec", default="f64", choices=["f64", "f32", "bf16x3"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chain", default="fused", choices=["fused", "4stage"],
                    help="LocalOp::product as run by the host layer (MPO pair fused) or the 4-stage chain")
    ap.add_argument("--no-qn", action="store_true", help="skip the block-sparse (31 Sz sectors) H*psi leg")
    ap.add_argument("--dmrg", default="100:full",
                    help="QN DMRG leg: N:maxdim1,maxdim2,... ('full' = the 10-sweep ramp to 2048); '' skips it")
    ap.add_argument("--dmrg-forms", default="f64,f32,bf16x3,bf16x3+f64last",
                    help="H*psi forms of the DMRG leg ('+f64last': the last sweep in FP64)")
    ap.add_argument("--hubbard", default="8x4:" + ",".join(str(m) for m in HUBBARD_MAXDIM),
                    help="Hubbard cylinder leg: NxxNy:maxdim1,maxdim2,... ('' skips it)")
    ap.add_argument("--hubbard-forms", default="f64,f32,bf16x3+f64last",
                    help="H*psi forms of the Hubbard leg ('+f64last': the last sweep in FP64)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="calibration only: time a plain device copy of this many GiB (for PMC byte counters)")
    args = ap.parse_args()
    if args.copy_gib > 0:
        return copy_calibration(args.copy_gib)
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)

    import torch

    from itensor_amd.itx import Context

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    ctx = Context(0, stream.cuda_stream)
    ctx.set_autotune(True)  # tile shapes measured per size class at plan creation (untimed)

    chi = args.chi
    rng = np.random.default_rng(SEED)
    psi = rng.standard_normal((chi, 2, 2, chi))
    L = rng.standard_normal((chi, 5, chi))
    R = rng.standard_normal((chi, 5, chi))
    W = heisenberg_w()

    def put(a):
        return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(dev)

    d_psi, d_L, d_R, d_W = put(psi), put(L), put(R), put(W)
    if args.chain == "fused":
        W12 = np.einsum("aijb,bklc->aijklc", W, W)  # W1*W2 over h1: (h0,s1,s1',s2,s2',h2)
        dims, labels = fused_dims(chi), FUSED
        ops = [d_L, put(W12), d_R]
    else:
        dims, labels = stage_dims(chi), STAGES
        ops = [d_L, d_W, d_W, d_R]
    nst = len(dims)
    bufs = [torch.empty(int(np.prod(dc)), dtype=torch.float64, device=dev) for (_, _, dc) in dims]
    inputs = [d_psi] + bufs[:-1]

    def run_forms(prec):
        plans = [ctx.plan_dense(da, la, db, lb, dc, lc, prec)
                 for (da, db, dc), (la, lb, lc) in zip(dims, labels)]
        return plans

    def hpsi(plans):
        for p, a, b, c in zip(plans, inputs, ops, bufs):
            p.execute(a.data_ptr(), b.data_ptr(), c.data_ptr())

    results = {}
    flops = hpsi_flops(chi)
    for prec in ([args.prec] + [p for p in ("f64", "f32", "bf16x3") if p != args.prec]):
        plans = run_forms(prec)
        for _ in range(args.warmup):
            hpsi(plans)
        torch.cuda.synchronize()
        # whole-step timing (wall clock around K steps)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            hpsi(plans)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # per-stage HIP-event timing on the stream the kernels run on
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(nst + 1)] for _ in range(args.steps)]
        for s in range(args.steps):
            for i, (p, a, b, c) in enumerate(zip(plans, inputs, ops, bufs)):
                ev[s][i].record(stream)
                p.execute(a.data_ptr(), b.data_ptr(), c.data_ptr())
            ev[s][nst].record(stream)
        torch.cuda.synchronize()
        stage_ms = [float(np.mean([ev[s][i].elapsed_time(ev[s][i + 1]) for s in range(args.steps)]))
                    for i in range(nst)]
        out = bufs[-1].detach().cpu().numpy().reshape((chi, 2, 2, chi), order="F")
        results[prec] = {"wall_s": wall, "stage_ms": stage_ms, "out": out,
                         "stage_flops": [p.flops for p in plans]}
        for p in plans:
            p.close()

    main_r = results[args.prec]
    ms = main_r["wall_s"] / args.steps * 1e3
    value = flops / (ms * 1e-3) / 1e12

    # dominant kernel: the stage with the largest average duration
    st = main_r["stage_ms"]
    dom = int(np.argmax(st))
    dom_flops = main_r["stage_flops"][dom]
    achieved = dom_flops / (st[dom] * 1e-3) / 1e12

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.prec,
        "data": "synthetic (psi, L, R iid N(0,1), seed %d; Heisenberg W)" % SEED,
        "config": {"workload": "dense two-site Hpsi, chi=%d, d=2, k=5 (Heisenberg MPO)" % chi,
                   "chi": chi, "flops_per_step": flops, "chain": args.chain, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK[args.prec],
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK[args.prec], 4),
                     "traffic": measured_traffic(chi, args.prec, ["S1", "W12", "S4"][dom] if nst == 3 else "none"),
                     "traffic_source": "profiles/pmc_dense_hpsi_chi1024.md (rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                                       "dispatch, scaled by a 1 GiB copy: x2.0 / x1.0)",
                     "kernel": "k_gemm (stage %d of %d of H*psi)" % (dom + 1, nst),
                     "flops_per_launch": dom_flops, "avg_launch_ms": round(st[dom], 4)},
        "stages_ms": {args.prec: [round(x, 4) for x in st]},
        "forms": {},
    }
    for prec, r in results.items():
        line["stages_ms"][prec] = [round(x, 4) for x in r["stage_ms"]]
        line["forms"][prec] = {"tflops": round(flops / (r["wall_s"] / args.steps) / 1e12, 3),
                               "ms_per_hpsi": round(r["wall_s"] / args.steps * 1e3, 4)}

    if not args.no_qn:
        q = qn_hpsi(max(5, args.steps), args.warmup)
        line["qn_hpsi"] = {"workload": "block-sparse two-site Hpsi, chi=1024 over 31 Sz sectors "
                                       "(max 120, min 1), site {+1,-1}, MPO link {0:3,-2:1,+2:1}", **q}

    if args.dmrg:
        n_s, md_s = args.dmrg.split(":")
        mds = DMRG_MAXDIM if md_s == "full" else [int(x) for x in md_s.split(",")]
        dm = {"workload": "QN Heisenberg N=%s open chain (total Sz 0, Neel start), maxdim %s, cutoff 1e-10, "
                          "niter 2; H*psi in the named form, everything else FP64" % (n_s, mds)}
        for form in args.dmrg_forms.split(","):
            base, _, last = form.partition("+")
            dm[form] = dmrg_leg(int(n_s), mds, {"f64": 0, "f32": 1, "bf16x3": 2}[base], 2,
                                1 if last == "f64last" else 0)
            print("dmrg leg %s done in %.1f s" % (form, dm[form]["wall_s"]), file=sys.stderr, flush=True)
        if "f64" in dm:
            for form in args.dmrg_forms.split(","):
                if form != "f64":
                    dm["dE_%s_vs_f64" % form] = dm[form]["energy"] - dm["f64"]["energy"]
                    dm["sweep_speedup_%s_vs_f64" % form] = round(dm["f64"]["wall_s"] / dm[form]["wall_s"], 3)
        line["dmrg"] = dm

    if args.hubbard:
        geo, md_s = args.hubbard.split(":")
        nx, ny = (int(x) for x in geo.split("x"))
        mds = [int(x) for x in md_s.split(",")]
        hb = {"workload": "Hubbard %dx%d YPeriodic cylinder (squareLattice order), t=1, U=4, half filling, "
                          "(Nf, Sz) QNs, MPO from the host AutoMPO; maxdim %s, cutoff 1e-9, niter 2, noise 1e-6/1e-7 "
                          "on the first two sweeps; H*psi in the named form, everything else FP64" % (nx, ny, mds)}
        for form in args.hubbard_forms.split(","):
            base, _, last = form.partition("+")
            hb[form] = hubbard_leg(nx, ny, mds, {"f64": 0, "f32": 1, "bf16x3": 2}[base], 1 if last == "f64last" else 0)
            print("hubbard leg %s done in %.1f s" % (form, hb[form]["wall_s"]), file=sys.stderr, flush=True)
        if "f64" in hb:
            for form in args.hubbard_forms.split(","):
                if form != "f64":
                    hb["dE_%s_vs_f64" % form] = hb[form]["energy"] - hb["f64"]["energy"]
                    hb["sweep_speedup_%s_vs_f64" % form] = round(hb["f64"]["wall_s"] / hb[form]["wall_s"], 3)
        line["hubbard"] = hb

    if not args.no_cpu:
        ref, t_cpu, threads, reps = cpu_baseline(psi, L, W, R)
        line["cpu_baseline"] = {"value": round(flops / t_cpu / 1e12, 5), "unit": "TFLOP/s",
                                "cores": threads, "kind": "port",
                                "sample": "oracle hpsi_dense (numpy/OpenBLAS) on the same chi=%d inputs, "
                                          "median of %d after 1 warm-up, %.3f s per H*psi" % (chi, reps, t_cpu)}
        nref = np.max(np.abs(ref))
        for prec, r in results.items():
            line["forms"][prec]["rel_err_vs_fp64_cpu"] = float(np.max(np.abs(r["out"] - ref)) / nref)
            line["forms"][prec]["rel_l2_err_vs_fp64_cpu"] = float(np.linalg.norm(r["out"] - ref) / np.linalg.norm(ref))
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
