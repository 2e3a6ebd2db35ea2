"""The measured workloads at their stated sizes, against the oracle.

* 4 x 4 periodic, t = 1, U = 8, 7 + 7 (D = 130,873,600): the 300-step FP64
  Lanczos ground energy from the benchmark's seeded vector against the oracle's
  OpenMP product-order Lanczos from the same vector (lanczos.h:130-175), to
  1e-10 relative.
* KPM density of states on the 1D ring L = 14, 7 + 7, U = 4 (D = 11,778,624):
  the Lanczos bounds against the oracle's Ritz values from the same start
  vector (1e-10 relative); the first 64 FP64 moments of one random-phase vector
  against the oracle's Chebyshev recurrence over its product-order H.v with one
  direct overlap per order (1e-11 absolute; kpm.h:212-229); the full 1024
  moments x 16 vectors in FP32 against FP64, with the moment error (bar 1e-6
  absolute) and the relative L1 error of the Jackson-damped DOS (bar 1e-6)
  stated and asserted.
* The serving mix at its stated lattices (1D L = 8..14, 2D 3 x 3 and 4 x 3),
  256-moment KPM and Krylov evolution up to t = 5 (dt = 0.05, k = 30), at a
  reduced request count: every served ground energy within 1e-8 of the oracle's
  Lanczos from the same seeded vector (D <= 1.2e6), every KPM request's bounds
  within 1e-8 of the oracle's Ritz values and its FP32 moments within 1e-5 of the
  oracle's FP64 moments on the same phase vectors (D <= 1.2e6), every Krylov
  request's overlap <psi0|psi(t)> and norm within 1e-9 of the oracle's stepped
  evolution (D <= 2e5); and every result is the same alone as in the mix.
"""
import os
from dataclasses import replace

import numpy as np
import pytest

import manybody_amd as mb
from manybody_amd.workloads import SERVE_LATTICES, serving_requests
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def ring_op(L, nu, nd, t, U):
    return mb.HubbardModel(t, U, L).operator(mb.FermionBasis.sector(L, nu, nd))


def ritz_extremes(alphas, betas):
    """Lowest and highest Ritz value by the oracle's Sturm bisection."""
    w = po.tridiag_eigvals(np.asarray(alphas), np.asarray(betas))
    return float(w.min()), float(w.max())


def device_ritz_extremes(dec):
    """Lowest and highest Ritz value by the library's tridiagonal solver."""
    w, _ = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)
    return float(np.min(w)), float(np.max(w))


# ------------------------------------------------------------------ 4 x 4, 7 + 7, U = 8

@pytest.mark.timeout(900)
def test_ground_energy_4x4_7p7_u8_300_steps_matches_oracle(cuda):
    """The benchmark's energy: seed 20260101, 300 steps, lowest Ritz value."""
    sizes, nu, nd, t, U = [4, 4], 7, 7, 1.0, 8.0
    op = mb.HubbardModel2D(t, U, 4, 4).operator(mb.FermionBasis.sector(16, nu, nd))
    x = op.gaussian(seed=20260101)
    dec = mb.lanczos_pass_one(op, x, 300)
    assert dec.steps_taken == 300
    e_dev = device_ritz_extremes(dec)[0]
    xh = mb.from_planar(x)
    del x, op
    P = po.ProductOperator(sizes, nu, nd, t, U, threads=THREADS)
    po.lib().orc_set_progress(25)  # a line every 25 host steps (minutes in all)
    try:
        al, be, _ = P.lanczos(xh, 300)
    finally:
        po.lib().orc_set_progress(0)
    e_orc = ritz_extremes(al, be)[0]
    print(f"4x4 7+7 U=8 300-step E0: device {e_dev:.14f} oracle {e_orc:.14f} "
          f"rel {abs(e_dev - e_orc) / abs(e_orc):.2e}")
    assert abs(e_dev - e_orc) / abs(e_orc) < 1e-10


# ------------------------------------------------------------------ KPM, L = 14, 7 + 7, U = 4

L14 = (14, 7, 7, 1.0, 4.0)


@pytest.fixture(scope="module")
def l14():
    L, nu, nd, t, U = L14
    return ring_op(L, nu, nd, t, U), po.ProductOperator([L], nu, nd, t, U, threads=THREADS)


def l14_bounds(op, k=100):
    b = op.gaussian(11)
    dec = mb.lanczos_pass_one(op, b, k)
    return b, dec


def test_kpm_l14_lanczos_bounds_match_oracle(cuda, l14):
    op, P = l14
    b, dec = l14_bounds(op)
    lo, hi = device_ritz_extremes(dec)
    al, be, _ = P.lanczos(mb.from_planar(b), 100)
    lo_o, hi_o = ritz_extremes(al, be)
    print(f"L=14 bounds: device [{lo:.14f}, {hi:.14f}] oracle [{lo_o:.14f}, {hi_o:.14f}]")
    assert abs(lo - lo_o) <= 1e-10 * abs(lo_o)
    assert abs(hi - hi_o) <= 1e-10 * abs(hi_o)


def test_kpm_l14_fp64_moments_match_oracle_recurrence(cuda, l14):
    op, P = l14
    b, dec = l14_bounds(op)
    lo, hi = device_ritz_extremes(dec)
    a, c = mb.rescaling_from_bounds(lo, hi, 1.0 / 0.99 - 1.0)
    M, seed = 64, 7
    mu = np.array(mb.kpm_dos_moments(op, a, c, M, 1, seed, fp32=False))
    r = mb.from_planar(op.random_phase(seed, 1, fp32=False))
    assert np.allclose(np.abs(r), 1.0, atol=1e-15)
    ref = P.kpm_moments(a, c, M, r)
    err = np.abs(mu - ref).max()
    print(f"L=14 first {M} FP64 moments: max |d mu| vs oracle {err:.3e}")
    assert err < 1e-11
    assert mu[0] == pytest.approx(1.0, abs=1e-14)


def test_kpm_l14_fp32_full_run_error_is_bounded(cuda, l14):
    op, _ = l14
    _, dec = l14_bounds(op)
    lo, hi = device_ritz_extremes(dec)
    a, c = mb.rescaling_from_bounds(lo, hi, 1.0 / 0.99 - 1.0)
    M, R, seed = 1024, 16, 7
    mu32 = np.array(mb.kpm_dos_moments(op, a, c, M, R, seed, fp32=True))
    mu64 = np.array(mb.kpm_dos_moments(op, a, c, M, R, seed, fp32=False))
    g = np.array(mb.jackson_kernel(M - 1))
    xs = np.linspace(-0.99, 0.99, 4001)
    Tn = np.cos(np.outer(np.arange(M), np.arccos(xs)))

    def dos(m):
        cn = g * m
        cn[1:] *= 2.0
        return (cn @ Tn) / (np.pi * np.sqrt(1.0 - xs ** 2))

    r32, r64 = dos(mu32), dos(mu64)
    l1 = float(np.trapezoid(np.abs(r32 - r64), xs) / np.trapezoid(np.abs(r64), xs))
    dmu = float(np.abs(mu32 - mu64).max())
    print(f"L=14 1024 x 16 FP32 vs FP64: max |d mu| {dmu:.3e}, DOS relative L1 {l1:.3e}")
    assert dmu < 1e-6
    assert l1 < 1e-6
    # the DOS integrates to one over the rescaled band
    assert np.trapezoid(r64, xs) == pytest.approx(1.0, abs=2e-2)


# ------------------------------------------------------------------ serving mix

def mix(n=24, seed=31, rate=200.0):
    """The serving workload's generator at its own lattices, moments and times."""
    return serving_requests(n, rate=rate, seed=seed, lattices=SERVE_LATTICES, ground_steps=150,
                            kpm_moments=256, kpm_nvec=4, kpm_bound_steps=40, t_max=5.0)


def product_of(req):
    return po.ProductOperator(list(req.sizes), req.n_up, req.n_dn, req.t, req.u,
                              threads=THREADS)


def op_of(req):
    L = int(np.prod(req.sizes))
    basis = mb.FermionBasis.sector(L, req.n_up, req.n_dn)
    if len(req.sizes) == 1:
        model = mb.HubbardModel(req.t, req.u, req.sizes[0])
    else:
        model = mb.HubbardModel2D(req.t, req.u, req.sizes[0], req.sizes[1])
    return model.operator(basis)


def check_against_oracle(req, got):
    """Returns True when the request was checked (small enough for the oracle)."""
    op = op_of(req)
    D = op.dimension()
    assert got["dim"] == D
    if req.kind == mb.REQ_EVOLVE and D > 200_000:
        return False
    if D > 1_200_000:
        return False
    P = product_of(req)
    b = mb.from_planar(op.gaussian(req.seed))
    if req.kind == mb.REQ_GROUND:
        al, be, _ = P.lanczos(b, req.steps)
        e = ritz_extremes(al, be)[0]
        assert abs(got["e0"] - e) <= 1e-8, (req, got["e0"], e)
    elif req.kind == mb.REQ_KPM:
        al, be, _ = P.lanczos(b, req.steps)
        lo, hi = ritz_extremes(al, be)
        margin = 0.01 * (hi - lo) + 1e-12  # the server's widening of the Ritz range
        assert abs(got["e_min"] - (lo - margin)) <= 1e-8, req
        assert abs(got["e_max"] - (hi + margin)) <= 1e-8, req
        a, c = mb.rescaling_from_bounds(got["e_min"], got["e_max"], 1.0 / 0.99 - 1.0)
        vecs = np.stack([mb.from_planar(op.random_phase(req.seed, req.nvec, fp32=False)
                                        [2 * r:2 * r + 2].contiguous())
                         for r in range(req.nvec)])
        ref = P.kpm_moments(a, c, req.moments, vecs)
        assert len(got["moments"]) == req.moments
        assert np.abs(got["moments"] - ref).max() <= 1e-5, req
    else:
        out = P.time_evolve_steps(b, 0.0, req.t_final, req.dt, req.steps)
        n0 = np.vdot(b, b).real
        ov = np.vdot(b, out) / n0
        assert abs(got["overlap_re"] - ov.real) <= 1e-9, (req, got["overlap_re"], ov.real)
        assert abs(got["overlap_im"] - ov.imag) <= 1e-9, (req, got["overlap_im"], ov.imag)
        assert abs(got["norm"] - np.linalg.norm(out) / np.sqrt(n0)) <= 1e-9, req
    return True


@pytest.mark.timeout(900)
def test_serving_mix_matches_oracle_and_alone(cuda):
    reqs = mix()
    kinds = np.bincount([r.kind for r in reqs], minlength=3)
    assert kinds.min() > 0, kinds  # every request kind is present
    assert max(r.t_final for r in reqs if r.kind == mb.REQ_EVOLVE) > 1.0
    with mb.Server(4, 16 << 30, lane_bytes=4 << 30) as S:
        ids = [S.submit(r) for r in reqs]
        mixed = [S.wait(i) for i in ids]
    for r, m in zip(reqs, mixed):
        assert m["status"] == 0, (r, m["error"])
    checked = [check_against_oracle(r, m) for r, m in zip(reqs, mixed)]
    per_kind = np.bincount([r.kind for r, c in zip(reqs, checked) if c], minlength=3)
    print(f"serving mix: {len(reqs)} requests {kinds.tolist()} by kind, oracle-checked "
          f"{per_kind.tolist()}, largest D {max(m['dim'] for m in mixed)}")
    assert per_kind.min() > 0
    # alone: one slot, no lane, each request on its own
    with mb.Server(1, 16 << 30) as S:
        for r, m in zip(reqs, mixed):
            alone = S.wait(S.submit(replace(r, arrival=-1.0)))
            assert alone["status"] == 0, alone["error"]
            if r.kind == mb.REQ_GROUND:
                # a lane member sums its tiles in another order than the two-pass H.v
                assert abs(alone["e0"] - m["e0"]) <= 1e-10 * max(1.0, abs(m["e0"])), r
            elif r.kind == mb.REQ_KPM:
                assert np.abs(alone["moments"] - m["moments"]).max() <= 1e-12, r
            else:
                for key in ("overlap_re", "overlap_im", "norm"):
                    assert abs(alone[key] - m[key]) <= 1e-12 * max(1.0, abs(m[key])), (r, key)
