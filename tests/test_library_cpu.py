"""CPU-side checks of libmanybody_hip.so: it loads, exports every symbol the C
header declares, and its host-only entry points (basis ranks, the tridiagonal
eigensolver, Jackson kernel, rescaling, argument validation) agree with the
oracle.  No GPU compute is called here."""
import re

import numpy as np
import pytest

import manybody_amd as mb
from oracle import pyoracle as po


def header_symbols():
    with open(mb.HEADER_PATH) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = mb.lib()
    syms = header_symbols()
    assert len(syms) >= 29
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(mb._SIGS), "ctypes signatures must cover the header"


def test_header_declares_every_exported_symbol():
    """The reverse: no mb_ symbol leaves the library undeclared (nm -D)."""
    import shutil
    import subprocess
    if shutil.which("nm") is None:
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-D", "--defined-only", mb.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.split()[-1].startswith("mb_")}
    assert exported == set(header_symbols())


def test_version():
    assert b"gfx950" in mb.lib().mb_version()


@pytest.mark.parametrize("L,nu,nd", [(2, 1, 1), (4, 2, 2), (6, 3, 2), (7, 0, 3), (8, 4, 4), (5, 5, 1),
                                     (9, 4, 3), (1, 1, 1)])
def test_basis_state_and_index_match_dfs(L, nu, nd):
    B = mb.FermionBasis.sector(L, nu, nd)
    up, dn = po.basis_states(L, nu, nd)
    assert B.size() == len(up)
    for i in range(len(up)):
        assert B.state(i) == (int(up[i]), int(dn[i]))
        assert B.index_of(int(up[i]), int(dn[i])) == i


def test_product_index_matches_oracle_permutation():
    for sizes, nu, nd in [([8], 4, 4), ([3, 3], 4, 3), ([10], 5, 5)]:
        L = int(np.prod(sizes))
        B = mb.FermionBasis.sector(L, nu, nd)
        assert np.array_equal(B.product_index(), po.product_permutation(sizes, nu, nd))


def test_c2_rank_subsample_matches_dfs():
    # 4 x 4, 7 + 7: D = 130,873,600; compare a strided subsample of the DFS order
    up, dn = po.basis_states(16, 7, 7)
    B = mb.FermionBasis.sector(16, 7, 7)
    assert B.size() == len(up) == 130873600
    for i in list(range(0, len(up), 9_999_991)) + [len(up) - 1]:
        assert B.state(i) == (int(up[i]), int(dn[i]))
        assert B.index_of(int(up[i]), int(dn[i])) == i


def test_basis_errors_follow_reference_contracts():
    with pytest.raises(ValueError):
        mb.FermionBasis.with_fixed_particle_number_and_spin(4, 3, 0)  # parity assert
    with pytest.raises(ValueError):
        mb.FermionBasis.with_fixed_particle_number_and_spin(4, 2, 4)  # |Sz| > N
    with pytest.raises(ValueError):
        mb.FermionBasis.with_fixed_particle_number_and_spin(3, 8, 0)  # N > 2L
    B = mb.FermionBasis.sector(4, 2, 1)
    with pytest.raises(IndexError):
        B.state(B.size())
    with pytest.raises(IndexError):
        B.index_of(0b0111, 0b0001)  # wrong up count: not in the basis
    assert not B.contains(0b1000, 0b0001) and B.contains(0b1001, 0b0001)


def test_tridiag_eig_matches_numpy_and_toeplitz():
    rng = np.random.default_rng(11)
    for m in [1, 2, 5, 30, 200]:
        a = rng.standard_normal(m)
        b = rng.standard_normal(max(m - 1, 0))
        T = np.diag(a) + (np.diag(b, 1) + np.diag(b, -1) if m > 1 else 0)
        w, V = mb.tridiag_eig(a, b)
        assert np.abs(w - np.linalg.eigvalsh(T)).max() < 1e-12 * max(1, np.abs(w).max())
        assert np.abs(T @ V - V * w).max() < 1e-11
        assert np.abs(V.T @ V - np.eye(m)).max() < 1e-12
    n = 300
    w, _ = mb.tridiag_eig(np.full(n, 2.0), np.full(n - 1, -1.0))
    exact = np.sort(2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
    assert np.abs(w - exact).max() < 1e-12


def test_tridiag_eig_agrees_with_oracle_bisection():
    rng = np.random.default_rng(12)
    a, b = rng.standard_normal(150), np.abs(rng.standard_normal(149))
    w, _ = mb.tridiag_eig(a, b, vectors=False)
    assert np.abs(w - po.tridiag_eigvals(a, b)).max() < 1e-12


def test_jackson_kernel_and_rescaling_match_oracle():
    for N in [1, 50, 1023]:
        assert np.array_equal(np.array(mb.jackson_kernel(N)), po.jackson_kernel(N))
    a, b = mb.rescaling_from_bounds(-3.0, 5.0, 0.01)
    assert a == (5.0 - -3.0) / 2.0 * 1.01 and b == 1.0


def test_request_footprint_covers_vectors_and_rejects_invalid():
    """Admission sizes (host arithmetic only): ground = start vector + three
    Lanczos vectors; KPM = start vector + max(Lanczos, three FP32 Chebyshev
    plane sets); evolution = start + output + the k + 1 stored Krylov vectors."""
    D = 3432 ** 2  # L = 14, 7 + 7
    g = mb.Request(mb.REQ_GROUND, (14,), 7, 7, u=4.0, steps=150).footprint()
    assert 64 * D <= g <= 64 * D + (64 << 20)
    k = mb.Request(mb.REQ_KPM, (14,), 7, 7, steps=40, moments=256, nvec=4).footprint()
    assert 16 * D + 3 * 8 * 4 * D <= k <= 16 * D + 3 * 8 * 4 * D + (64 << 20)
    e = mb.Request(mb.REQ_EVOLVE, (14,), 7, 7, steps=30, t_final=1.0).footprint()
    assert 16 * D * (2 + 31) <= e <= 16 * D * (2 + 31) + (64 << 20)
    assert mb.Request(mb.REQ_GROUND, (40,), 1, 1).footprint() == -1
    assert mb.Request(mb.REQ_GROUND, (8,), 9, 1).footprint() == -1
    assert mb.Request(7, (8,), 2, 2).footprint() == -1


def test_optical_conductivity_matches_reference_golden():
    """compute_optical_conductivity on the host (optical_conductivity.cpp:8-80)
    against the compiled reference's outputs for four correlators."""
    import json
    import os
    gold_dir = os.path.join(os.path.dirname(__file__), "golden")
    with open(os.path.join(gold_dir, "manifest.json")) as f:
        cases = [c for c in json.load(f)["conductivity"] if c["kind"] == "sigma"]
    gold = np.load(os.path.join(gold_dir, "ref_conductivity.npz"))
    assert len(cases) == 4
    for c in cases:
        r = mb.compute_optical_conductivity(gold[c["name"] + "__corr"], c["dt"], c["beta"],
                                            c["volume"])
        w = gold[c["name"] + "__omega"]
        assert r.frequencies.shape == w.shape
        assert np.abs(r.frequencies - w).max() <= 1e-12 * w.max()
        ref = gold[c["name"] + "__sigma"]
        assert np.abs(r.sigma - ref).max() <= 1e-12 * np.abs(ref).max()
        assert abs(r.sum_rule - c["sum_rule"]) <= 1e-12 * max(abs(c["sum_rule"]), 1e-300)


def test_optical_conductivity_rejects_invalid_arguments():
    with pytest.raises(ValueError):
        mb.compute_optical_conductivity(np.ones(1), 0.1, 1.0, 1.0)
    with pytest.raises(ValueError):
        mb.compute_optical_conductivity(np.ones(8), 0.0, 1.0, 1.0)
    with pytest.raises(ValueError):
        mb.compute_optical_conductivity(np.ones(8), 0.1, 1.0, 0.0)


def _gold(name):
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", name))


def _manifest():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "manifest.json")) as f:
        return json.load(f)


def _oracle_matrix(L, states, expr):
    terms = [(c, [(o >> 15, (o >> 14) & 1, o & 0x3FFF) for o in k]) for k, c in expr.terms.items()]
    return po.expression_matrix(L, states, terms)


def test_model_expressions_match_reference_golden():
    """HubbardModel{,2D,3D}.hamiltonian() and HubbardModelMomentum.hamiltonian()
    built by the Python mirror give the compiled reference's matrices (28
    real-space and the momentum-space golden sectors)."""
    H = _gold("ref_hamiltonian.npz")
    for e in _manifest()["hamiltonian"]:
        dims = e["dims"]
        model = (mb.HubbardModel(e["t"], e["U"], dims[0]) if len(dims) == 1 else
                 mb.HubbardModel2D(e["t"], e["U"], *dims) if len(dims) == 2 else
                 mb.HubbardModel3D(e["t"], e["U"], *dims))
        L = int(np.prod(dims))
        up, dn = po.basis_states(L, e["n_up"], e["n_dn"])
        M = _oracle_matrix(L, list(zip(up, dn)), model.hamiltonian())
        R = np.zeros_like(M)
        ij = H[e["name"] + "__ij"]
        R[ij[:, 0], ij[:, 1]] = H[e["name"] + "__val"]
        assert np.abs(M - R).max() <= 1e-14, e["name"]
    G = _gold("ref_momentum.npz")
    n = 0
    for e in _manifest()["momentum"]:
        if e["kind"] != "hamiltonian":
            continue
        dims = e["dims"]
        st = G[e["name"] + "__states"]
        M = _oracle_matrix(int(np.prod(dims)), st,
                           mb.HubbardModelMomentum(e["t"], e["U"], dims).hamiltonian())
        R = np.zeros_like(M)
        ij = G[e["name"] + "__ij"]
        R[ij[:, 0], ij[:, 1]] = G[e["name"] + "__val"]
        assert np.abs(M - R).max() <= 1e-12, e["name"]
        n += 1
    assert n >= 10


def test_fermion_expression_algebra_follows_add_to():
    e = mb.hopping(0.5 + 0.25j, 0, 2, mb.Spin.Down)
    assert len(e) == 2 and e.adjoint().terms == e.terms  # hopping is Hermitian
    e += mb.density(mb.Spin.Up, 1) * 1e-14            # below 1000 eps: dropped
    assert len(e) == 2
    n = mb.density(mb.Spin.Up, 1)
    z = n * 1.5 - n * 1.5                              # cancels: erased
    assert len(z) == 0
    dd = mb.density_density(mb.Spin.Up, 0, mb.Spin.Down, 0)
    assert list(dd.terms) == [(0x0000, 0x8000, 0x4000, 0xC000)]
    two = mb.two_body(0, 1, 1, 2, 1, 3, 0, 4)
    assert list(two.terms) == [(0x0001, 0x4002, 0xC003, 0x8004)]
    prod = (mb.creation(0, 0) + 1.0) * mb.annihilation(1, 3)
    assert prod.terms == {(0x0000, 0xC003): 1.0, (0xC003,): 1.0}
    assert (2.0 * prod).terms[(0xC003,)] == 2.0 and len(prod * 0.0) == 0


def test_sparse_matrix_cereal_matches_reference_golden(tmp_path):
    """SparseMatrix.save writes the bytes the vendored cereal wrote for the same
    matrices; load reads them back."""
    gold = _gold("ref_expression.npz")
    n = 0
    for c in _manifest()["expression"]:
        if c["kind"] != "cereal":
            continue
        rows, cols, vals = (gold[c["name"] + s] for s in ("__rows", "__cols", "__vals"))
        sp = mb.SparseMatrix.from_entries((c["n_rows"], c["n_cols"]), rows, cols, vals)
        path = tmp_path / (c["name"] + ".bin")
        sp.save(str(path))
        assert path.read_bytes() == gold[c["name"] + "__bytes"].tobytes(), c["name"]
        back = mb.SparseMatrix.load(str(path))
        r2, c2, v2 = back.entries()
        assert back.shape == (c["n_rows"], c["n_cols"]) and back.nnz == c["nnz"]
        assert np.array_equal(r2, rows) and np.array_equal(c2, cols) and np.array_equal(v2, vals)
        n += 1
    assert n == 3
    with pytest.raises(ValueError):  # duplicate entry
        mb.SparseMatrix.from_entries((2, 2), [0, 0], [1, 1], [1.0, 2.0])
    with pytest.raises(RuntimeError):
        mb.SparseMatrix.load(str(tmp_path / "missing.bin"))


def test_kpm_host_functions_match_oracle():
    """projector_moments, lorentz_kernel and jackson_kernel of the library against
    the oracle (closed forms), and |epsilon| > 1 rejected."""
    for M, eps in [(0, 0.3), (10, -0.8), (200, 0.0), (64, 1.0)]:
        assert np.allclose(mb.projector_moments(M, eps), po.projector_moments(M, eps),
                           rtol=0, atol=1e-15)
    assert np.allclose(mb.lorentz_kernel(50, 4.0), po.lorentz_kernel(50, 4.0), rtol=1e-14)
    with pytest.raises(ValueError):
        mb.projector_moments(10, 1.5)


def test_every_entry_point_rejects_null_arguments():
    """Every status-returning entry point, called with null handles and pointers
    and zero sizes, returns a status code (no crash); all but the empty
    tridiagonal problem report an error.  Run in a child process so that a
    crash fails this test instead of the session."""
    import subprocess
    import sys
    code = r'''
import ctypes as C, sys
sys.path.insert(0, %r)
import manybody_amd as mb
L = mb.lib()
for name in sorted(mb._SIGS):
    res, args = mb._SIGS[name]
    if res is not C.c_int:
        continue
    vals = [a(0) if a in (C.c_int, C.c_int64, C.c_uint64, C.c_double) else None for a in args]
    print(name, getattr(L, name)(*vals), flush=True)
''' % (str(__import__("os").path.dirname(mb.__path__[0])),)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    status = dict(line.split() for line in r.stdout.splitlines())
    assert len(status) > 60
    ok_with_nothing = {"mb_tridiag_eig"}  # m = 0: nothing to do
    for name, st in status.items():
        assert int(st) in range(6), name
        assert (int(st) == 0) == (name in ok_with_nothing), (name, st)


def test_fixed_n_basis_matches_reference_golden():
    """with_fixed_particle_number through the library (host side): states and
    indices of the compiled reference's lists for every L <= 6, N <= min(2L, 7)."""
    gold = _gold("ref_expression.npz")
    n = 0
    for c in _manifest()["expression"]:
        if c["kind"] != "nbasis":
            continue
        b = mb.FermionBasis.with_fixed_particle_number(c["L"], c["N"])
        st = gold[c["name"] + "__states"]
        assert b.size() == len(st)
        for i, (u, d) in enumerate(st):
            assert b.state(i) == (int(u), int(d))
            assert b.index_of(int(u), int(d)) == i
        n += 1
    assert n == 39
    with pytest.raises(ValueError):
        mb.FermionBasis.with_fixed_particle_number(3, 7)


def test_fixed_n_basis_reference_unit_cases():
    """Ports of test_basis.cpp:5-19: with_fixed_particle_number(1, 1) holds
    c+_up0 and c+_dn0 (in that order, the reference's enumeration), (1, 2) the
    one state c+_up0 c+_dn0."""
    b = mb.FermionBasis.with_fixed_particle_number(1, 1)
    assert b.size() == 2
    assert b.state(0) == (1, 0) and b.state(1) == (0, 1)
    b = mb.FermionBasis.with_fixed_particle_number(1, 2)
    assert b.size() == 1 and b.state(0) == (1, 1)


def test_momentum_current_expression_validation():
    """HubbardModelMomentum::current (hubbard_model_momentum.h:52-60): Q of the
    wrong dimension and an out-of-range direction are invalid arguments; J(Q)
    is sum_k v(k) c+_{k+Q} c_k for both spins."""
    m = mb.HubbardModelMomentum(1.0, 4.0, [4, 2])
    with pytest.raises(ValueError):
        m.current([1], 0)
    with pytest.raises(ValueError):
        m.current([1, 0], 2)
    j = m.current([1, 1], 0)
    assert len(j.terms) == 2 * 4  # kx in {1, 3} (v != 0) x ky in {0, 1} x 2 spins
    # k = (1, 0) -> k + Q = (2, 1), flat 2 + 4 * 1 = 6; v = 2 (2 pi / 4) sin(pi / 2) = pi
    assert abs(j.terms[(6, 0x8000 | 1)] - np.pi) < 1e-15
    assert abs(j.terms[(0x4000 | 6, 0xC000 | 1)] - np.pi) < 1e-15


def test_reference_rescaling_host_cases():
    """test_rescaling.cpp:9-92: rescale / inverse and from_bounds."""
    r = mb.Rescaling(2.5, 1.0)
    assert abs(r.inverse(r.rescale(3.5)) - 3.5) < 1e-14
    r = mb.Rescaling(3.0, 2.0)
    assert abs(r.rescale(2.0)) < 1e-14 and abs(r.rescale(5.0) - 1.0) < 1e-14
    assert abs(r.rescale(-1.0) + 1.0) < 1e-14
    assert abs(r.inverse(0.0) - 2.0) < 1e-14 and abs(r.inverse(1.0) - 5.0) < 1e-14
    assert abs(r.inverse(-1.0) + 1.0) < 1e-14
    r = mb.Rescaling.from_bounds(-2.0, 4.0, 0.0)
    assert abs(r.a - 3.0) < 1e-14 and abs(r.b - 1.0) < 1e-14
    r = mb.Rescaling.from_bounds(-2.0, 4.0, 0.1)
    assert abs(r.a - 3.3) < 1e-14 and abs(r.b - 1.0) < 1e-14


def test_host_only_use_does_not_import_torch():
    """On a machine without a GPU, loading the library and using its host-side
    entry points (bases, index lookups, KPM kernels, rescaling, the tridiagonal
    solver) imports no torch (with a GPU, torch's HIP runtime is brought up
    before the library loads).  Child process: nothing imported beforehand."""
    import subprocess
    import sys
    code = r'''
import sys
sys.path.insert(0, %r)
import manybody_amd as mb
mb.lib()
b = mb.FermionBasis.sector(6, 3, 3)
assert b.index_of(*b.state(7)) == 7
mb.jackson_kernel(8); mb.rescaling_from_bounds(-2.0, 3.0); mb.projector_moments(8, 0.1)
mb.tridiag_eig([1.0, 2.0], [0.5])
print("torch imported:", "torch" in sys.modules)
''' % (str(__import__("os").path.dirname(mb.__path__[0])),)
    if __import__("os").path.exists("/dev/kfd"):
        pytest.skip("a GPU is present: torch is initialised before the library loads")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "torch imported: False" in r.stdout


def test_host_sanitizer_driver_host_part():
    """tests/cpp/test_host_sanitize.cpp against the host-sanitized library
    (AddressSanitizer with leak detection, UndefinedBehaviorSanitizer): bases,
    lookups, error paths, host kernels, the sparse container and its cereal
    file, request footprints.  Any report fails the run (nonzero status)."""
    import os
    import subprocess
    binary = os.path.join(os.path.dirname(__file__), "cpp", "_build", "test_host_sanitize")
    if not os.path.exists(binary):
        pytest.skip("host-sanitizer driver not built (__graft_entry__.build())")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1",
               HIP_VISIBLE_DEVICES="")
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "HOST PASSED" in r.stdout
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
