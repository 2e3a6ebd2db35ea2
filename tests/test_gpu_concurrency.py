"""One operator shared by several host threads, each on its own stream (the
reference's OpenMP sample loops share one const operator with per-thread
vectors, hubbard_momentum_optical_conductivity_factorized.cpp:226-283).  H.v,
reference-order conversion, seeded fills and Lanczos run from four threads at
once and must give the serial results bit for bit (every kernel is
deterministic: fixed reduction trees, no atomics).  A fresh basis whose device
tables are first needed by the threads themselves checks the one-time upload.
"""
import threading

import numpy as np
import pytest
import torch

import manybody_amd as mb

pytestmark = pytest.mark.gpu

THREADS = 4


def run_threads(fn):
    out, errors = [None] * THREADS, []

    def body(k):
        try:
            s = torch.cuda.Stream()  # default priority
            with torch.cuda.stream(s):
                r = fn(k)
            s.synchronize()
            out[k] = r
        except BaseException as e:  # reported in the main thread
            errors.append(e)

    ts = [threading.Thread(target=body, args=(k,)) for k in range(THREADS)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts), "a worker thread did not finish"
    if errors:
        raise errors[0]
    return out


def work(op, k):
    x = op.gaussian(seed=100 + k)
    y = op.hv(x)
    for _ in range(4):  # repeated launches on the thread's stream
        y = op.hv(x, y)
    dec = mb.lanczos_pass_one(op, x, 25)
    ref = op.to_reference_order(y)
    return (y.cpu().numpy(), np.array(dec.alphas), np.array(dec.betas), ref.cpu().numpy())


def operators():
    real = mb.HubbardModel2D(1.0, 6.0, 3, 4).operator(mb.FermionBasis.sector(12, 3, 3))
    mbasis = mb.FermionBasis.with_fixed_particle_number_spin_momentum(9, 6, 0, [3, 3], [1, 2])
    mom = mb.HubbardMomentumFactorizedFixedK(mbasis, [3, 3], 1.0, 6.0, [1, 2])
    e = mb.HubbardModel(1.0, 4.0, 10).hamiltonian() + mb.hopping(0.3j, 0, 5, mb.Spin.Up)
    expr = mb.FermionExpressionOperator(mb.FermionBasis.sector(10, 4, 4), e)
    comp = mb.Shifted(real, 0.5) + 0.25 * real  # operator algebra: scratch per call
    return {"real": real, "momentum": mom, "expression": expr, "composite": comp}


@pytest.mark.parametrize("kind", ["real", "momentum", "expression", "composite"])
def test_concurrent_calls_match_serial(cuda, kind):
    op = operators()[kind]
    serial = [work(op, k) for k in range(THREADS)]
    torch.cuda.synchronize()
    conc = run_threads(lambda k: work(op, k))
    for k in range(THREADS):
        for a, b in zip(serial[k], conc[k]):
            assert np.array_equal(a, b), (kind, k)


def test_first_device_upload_from_threads(cuda):
    """The sector's device tables are uploaded by whichever thread needs them
    first; the others wait and use the same tables."""
    e = mb.HubbardModel(1.0, 4.0, 9).hamiltonian()
    want = None
    for trial in range(3):
        op = mb.FermionExpressionOperator(mb.FermionBasis.sector(9, 4, 3), e)
        got = run_threads(lambda k: op.gaussian(seed=7).cpu().numpy())
        for g in got[1:]:
            assert np.array_equal(g, got[0])
        if want is None:
            want = got[0]
        assert np.array_equal(got[0], want)


def test_library_first_then_torch_in_a_fresh_process(cuda):
    """A script that builds operators before touching torch.cuda still gets a
    device for torch (the package brings torch's HIP runtime up first); run in a
    child process, where nothing has initialised the GPU yet."""
    import os
    import subprocess
    import sys
    code = (
        "import manybody_amd as mb\n"
        "b = mb.FermionBasis.sector(6, 3, 3)\n"
        "op = mb.FermionExpressionOperator(b, mb.HubbardModel(1.0, 4.0, 6).hamiltonian())\n"
        "e = mb.find_min_eigenpair(op, 100, op.gaussian(1), want_vector=False).value\n"
        "print('ok', e)\n")
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=repo, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-1000:] + r.stderr[-2000:]
