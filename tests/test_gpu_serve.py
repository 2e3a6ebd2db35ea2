"""Serving loop (continuous admission of mixed requests).

* Every result equals the same request run alone, and the same computation
  through the library's direct entry points: ground energies, KPM moments and
  time-evolved overlaps to 1e-12 relative (every reduction in the library has a
  fixed order, so they agree bit for bit in practice; the tolerance is the
  stated bar).
* Admission respects the arena: at no admission instant do the footprints of
  the running requests exceed it; a request larger than the arena fails with
  MB_ERR_NOMEM without disturbing the others; no request uses more device
  memory than it was admitted with.
* Requests arriving over time are admitted after they arrive and overlap.
"""
from dataclasses import replace

import numpy as np
import pytest
import torch

import manybody_amd as mb
from manybody_amd.workloads import serving_requests

pytestmark = pytest.mark.gpu

ARENA = 2 << 30
SMALL_LATTICES = [(8,), (9,), (10,), (3, 3)]


def small_mix(n=24, seed=11, rate=400.0):
    return serving_requests(n, rate=rate, seed=seed, lattices=SMALL_LATTICES, ground_steps=80,
                            kpm_moments=64, kpm_nvec=2, kpm_bound_steps=30, t_max=0.5)


def close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(1.0, abs(a), abs(b))


def check_same(req, a, b):
    assert a["status"] == 0 and b["status"] == 0, (a["error"], b["error"])
    assert a["dim"] == b["dim"]
    if req.kind == mb.REQ_GROUND:
        assert close(a["e0"], b["e0"]), (req, a["e0"], b["e0"])
    elif req.kind == mb.REQ_KPM:
        assert close(a["e_min"], b["e_min"]) and close(a["e_max"], b["e_max"])
        assert np.abs(a["moments"] - b["moments"]).max() <= 1e-12, req
    else:
        assert close(a["overlap_re"], b["overlap_re"]) and close(a["overlap_im"], b["overlap_im"])
        assert close(a["norm"], b["norm"]), req


def op_of(req):
    L = int(np.prod(req.sizes))
    basis = mb.FermionBasis.sector(L, req.n_up, req.n_dn)
    if len(req.sizes) == 1:
        model = mb.HubbardModel(req.t, req.u, req.sizes[0])
    else:
        model = mb.HubbardModel2D(req.t, req.u, req.sizes[0], req.sizes[1])
    return model.operator(basis)


def direct(req):
    """The same computation through the library's one-at-a-time entry points."""
    op = op_of(req)
    b = op.gaussian(req.seed)
    if req.kind == mb.REQ_GROUND:
        return {"e0": mb.find_min_eigenpair(op, req.steps, b, want_vector=False).value}
    if req.kind == mb.REQ_KPM:
        dec = mb.lanczos_pass_one(op, b, req.steps)
        w, _ = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)
        lo, hi = w[0], w[-1]
        margin = 0.01 * (hi - lo) + 1e-12
        e_min, e_max = lo - margin, hi + margin
        a = (e_max - e_min) / 2.0 * (1.0 + (1.0 / 0.99 - 1.0))
        c = (e_max + e_min) / 2.0
        mu = mb.kpm_dos_moments(op, a, c, req.moments, req.nvec, req.seed, fp32=True)
        return {"e_min": e_min, "e_max": e_max, "moments": np.array(mu)}
    out = mb.time_evolve_state_steps(op, b, 0.0, req.t_final, req.dt,
                                      mb.EvolutionOptions(req.steps))
    z0 = mb.from_planar(b)
    z1 = mb.from_planar(out)
    n0 = np.vdot(z0, z0).real
    ov = np.vdot(z0, z1) / n0
    return {"overlap_re": ov.real, "overlap_im": ov.imag, "norm": np.linalg.norm(z1) / np.sqrt(n0)}


def test_serve_results_match_direct_calls(cuda):
    reqs = small_mix(12, seed=3)
    with mb.Server(4, ARENA) as S:
        res = [S.wait(S.submit(replace(r, arrival=-1.0))) for r in reqs]
    for r, got in zip(reqs, res):
        assert got["status"] == 0, got["error"]
        ref = direct(r)
        for key, val in ref.items():
            if key == "moments":
                assert np.abs(got["moments"] - val).max() <= 1e-12, r
            else:
                assert close(got[key], val, 1e-12), (r, key, got[key], val)


def test_serve_alone_equals_mixed(cuda):
    reqs = small_mix(32, seed=11)
    with mb.Server(8, ARENA) as S:
        ids = [S.submit(r) for r in reqs]
        mixed = [S.wait(i) for i in ids]
    # the mix did overlap: some request was admitted before an earlier one finished
    overlaps = sum(1 for i in range(1, len(mixed))
                   if any(m["admitted"] < mixed[i]["finished"] and m["finished"] > mixed[i]["admitted"]
                          for m in mixed[:i]))
    assert overlaps > 0
    with mb.Server(1, ARENA) as S:
        for r, m in zip(reqs, mixed):
            alone = S.wait(S.submit(replace(r, arrival=-1.0)))
            check_same(r, alone, m)


def test_serve_admission_respects_arena(cuda):
    reqs = [replace(r, arrival=-1.0) for r in small_mix(24, seed=7)]
    fps = [r.footprint() for r in reqs]
    arena = 3 * max(fps)
    with mb.Server(8, arena) as S:
        res = [S.wait(S.submit(r)) for r in reqs]
    for r, x in zip(reqs, res):
        assert x["status"] == 0, x["error"]
        assert 0 < x["peak"] <= x["footprint"], (r, x["peak"], x["footprint"])
        assert x["peak"] >= 16 * x["dim"]
    arena_aligned = arena // 256 * 256
    for x in res:  # memory held at each admission instant
        t = x["admitted"]
        held = sum(y["footprint"] for y in res if y["admitted"] <= t < y["finished"])
        assert held <= arena_aligned


def test_serve_rejects_request_larger_than_arena(cuda):
    small = [replace(r, arrival=-1.0) for r in small_mix(6, seed=9)]
    big = mb.Request(mb.REQ_GROUND, (12,), 6, 6, u=4.0, steps=40, seed=1)
    arena = max(r.footprint() for r in small) * 2
    assert big.footprint() > arena
    with mb.Server(4, arena) as S:
        ids = [S.submit(r) for r in small[:3]] + [S.submit(big)] + [S.submit(r) for r in small[3:]]
        res = [S.wait(i) for i in ids]
    assert res[3]["status"] == mb.MB_ERR_NOMEM and "arena" in res[3]["error"]
    assert all(x["status"] == 0 for k, x in enumerate(res) if k != 3)


def test_serve_poisson_arrivals_are_respected(cuda):
    reqs = small_mix(30, seed=21, rate=150.0)
    with mb.Server(6, ARENA) as S:
        ids = [S.submit(r) for r in reqs]
        res = [S.wait(i) for i in ids]
    for r, x in zip(reqs, res):
        assert x["status"] == 0, x["error"]
        assert x["arrival"] == pytest.approx(r.arrival)
        assert x["admitted"] >= x["arrival"] - 1e-6
        assert x["finished"] >= x["admitted"]


# ------------------------------------------------------------------ Lanczos lane

LANE = 1 << 30


def lane_mix(n=40, seed=17, rate=300.0):
    return serving_requests(n, rate=rate, seed=seed, lattices=[(8,), (10,), (12,), (3, 3), (4, 3)],
                            ground_steps=120, kpm_moments=64, kpm_nvec=2, kpm_bound_steps=30,
                            t_max=0.3)


def test_lane_ground_energies_match_direct_path(cuda):
    """Lane members run the batched kernels (tile sums in another order than the
    two-pass H.v), so energies agree with the one-at-a-time path to 1e-10
    relative, the energy bar, not bit for bit."""
    reqs = [r for r in lane_mix(30, seed=2) if r.kind == mb.REQ_GROUND][:12]
    with mb.Server(2, 4 << 30, lane_bytes=LANE) as S:
        res = [S.wait(S.submit(replace(r, arrival=-1.0))) for r in reqs]
    for r, x in zip(reqs, res):
        assert x["status"] == 0, x["error"]
        assert x["slot"] >= 2, "ground request did not run in the lane"
        e = direct(r)["e0"]
        assert abs(x["e0"] - e) <= 1e-10 * max(1.0, abs(e)), (r, x["e0"], e)


def test_lane_alone_equals_mixed(cuda):
    reqs = lane_mix(40, seed=17)
    with mb.Server(4, 6 << 30, lane_bytes=LANE) as S:
        ids = [S.submit(r) for r in reqs]
        mixed = [S.wait(i) for i in ids]
    in_lane = [m for r, m in zip(reqs, mixed) if r.kind == mb.REQ_GROUND and m["slot"] >= 4]
    assert len(in_lane) >= 10
    # members joined while others were in the lane
    joins = sum(1 for a in in_lane for b in in_lane
                if a is not b and b["admitted"] < a["admitted"] < b["finished"])
    assert joins > 0
    with mb.Server(1, 6 << 30, lane_bytes=LANE) as S:
        for r, m in zip(reqs, mixed):
            alone = S.wait(S.submit(replace(r, arrival=-1.0)))
            check_same(r, alone, m)
            if r.kind == mb.REQ_GROUND:
                assert alone["e0"] == m["e0"]  # same tiles, same order: bit for bit


def test_lane_breakdown_and_tiny_models(cuda):
    """D = 1 (breakdown at the first step) and D < steps finish with the exact
    lowest eigenvalue; they share the lane with a normal member."""
    reqs = [mb.Request(mb.REQ_GROUND, (8,), 0, 0, u=3.0, steps=50, seed=1),    # D = 1, E = 0
            mb.Request(mb.REQ_GROUND, (8,), 8, 8, u=3.0, steps=50, seed=2),    # D = 1, E = 8 U
            mb.Request(mb.REQ_GROUND, (4,), 1, 1, u=2.0, steps=40, seed=3),    # D = 16 < steps
            mb.Request(mb.REQ_GROUND, (10,), 5, 5, u=4.0, steps=100, seed=4)]
    with mb.Server(1, 2 << 30, lane_bytes=LANE) as S:
        res = [S.wait(S.submit(r)) for r in reqs]
    assert all(x["status"] == 0 for x in res), [x["error"] for x in res]
    assert res[0]["e0"] == 0.0
    assert abs(res[1]["e0"] - 24.0) < 1e-12
    e_exact = np.linalg.eigvalsh(op_dense(reqs[2]))[0]
    assert abs(res[2]["e0"] - e_exact) < 1e-10
    assert abs(res[3]["e0"] - direct(reqs[3])["e0"]) < 1e-10 * abs(res[3]["e0"])


def op_dense(req):
    op = op_of(req)
    D = op.dimension()
    cols = []
    for j in range(D):
        e = np.zeros(D, np.complex128)
        e[j] = 1.0
        cols.append(op.apply(torch.from_numpy(e).cuda()).cpu().numpy())
    return np.array(cols).T.real
