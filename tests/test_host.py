"""CPU tests of the host side: the C-ABI library loads and exports every declared symbol,
host-only entry points, problem set-up vs the oracle's restatement of am_topopt.cpp."""
import os
import re

import numpy as np
import pytest

import am_ml_suite_amd as A
from am_ml_suite_amd import _ffi
import oracle_ffi as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "amt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(amt_[a-z0-9_]+)\s*\(", src)) - {"amt_iter_cb"})


def test_library_exports_every_declared_symbol():
    L = _ffi.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_ffi.EXPORTS) == syms


def test_element_stiffness_matches_oracle():
    for E, nu, h in [(1.0, 0.3, 1.0), (113.8e9, 0.342, 0.004), (3.0, 0.1, 0.5)]:
        a = A.element_stiffness(E, nu, h)
        b = O.ke0(E, nu, h)
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


def test_element_stiffness_rejects_bad_material():
    with pytest.raises(ValueError):
        A.element_stiffness(1.0, 0.5, 1.0)


def test_create_fails_loudly_without_device_or_bad_config():
    cfg = A.SIMPConfig()
    with pytest.raises((ValueError, A.AmtError)):
        A.TopologyOptimizer(0, 1, 1, 1.0, A.MaterialProperties(), cfg)
    if not A.is_available():
        with pytest.raises(A.AmtError):
            A.TopologyOptimizer(2, 2, 2, 1.0, A.MaterialProperties(), cfg)


@pytest.mark.parametrize("res", [4.0, 12.0, 6.0])
def test_bracket_domain_matches_oracle(res):
    dom = A.bracket_domain(res)
    g, shape = O.bracket(res)
    assert dom.shape == shape
    assert np.array_equal(dom.slot_order(dom.grid), g)
    lc = A.bracket_load_case(dom)
    F, _ = O.bracket_forces(*shape, res)
    assert np.array_equal(lc.get_force_vector(), F)
    assert np.array_equal(lc.get_constrained_dofs(), O.bracket_fixed(*shape))


def test_cantilever_problem():
    p = A.cantilever_problem()
    assert p.n_dofs == 11253 and len(p.fixed) == 363
    assert np.nonzero(p.F)[0].tolist() == [11057] and p.F[11057] == -1.0


def test_bracket_like_problem_deterministic():
    a = A.bracket_like_problem()
    b = A.bracket_like_problem()
    assert a.n_dofs == 1350723 and a.n_elements == 432000
    assert np.array_equal(a.rho0, b.rho0) and np.array_equal(a.F, b.F)
    assert np.count_nonzero(a.F) > 0 and len(a.fixed) == 2 * 11 * 11 * 3
    assert (a.rho0 == 1.0).sum() == 2000 and (a.rho0 == 0.0).sum() > 0
    c = A.bracket_like_problem(seed=3)
    assert not np.array_equal(a.F, c.F)


def test_batch_problem_generator():
    """The dataset-generation set: deterministic, sizes in U{16..64}, one whole clamped face
    (all three DOFs of its nodes), 1-4 point loads of magnitude in [0.5, 1] on other faces,
    volume fraction in [0.2, 0.5], filter radius in [1.2, 2.5]."""
    ps = A.batch_problems()
    qs = A.batch_problems()
    assert len(ps) == 64
    for p, q in zip(ps, qs):
        assert (p.nx, p.ny, p.nz) == (q.nx, q.ny, q.nz)
        assert np.array_equal(p.F, q.F) and np.array_equal(p.fixed, q.fixed)
        assert all(16 <= d <= 64 for d in (p.nx, p.ny, p.nz))
        nn = np.array([p.nx + 1, p.ny + 1, p.nz + 1])
        nodes = np.unique(p.fixed // 3)
        assert p.fixed.size == 3 * nodes.size
        I, J, K = nodes // (nn[1] * nn[2]), (nodes // nn[2]) % nn[1], nodes % nn[2]
        on = [np.all(I == 0), np.all(I == p.nx), np.all(J == 0), np.all(J == p.ny),
              np.all(K == 0), np.all(K == p.nz)]
        face = on.index(True)
        assert nodes.size == np.prod(np.delete(nn, face // 2))
        loaded = np.unique(np.nonzero(p.F)[0] // 3)
        assert 1 <= loaded.size <= 4
        assert not np.any(np.isin(loaded, nodes))
        mags = np.linalg.norm(p.F.reshape(-1, 3)[loaded], axis=1)
        assert np.all(mags >= 0.5 - 1e-12) and np.all(mags <= 4.0)
        assert 0.2 <= p.params["volume_fraction"] <= 0.5
        assert 1.2 <= p.params["filter_radius"] <= 2.5
        assert np.all(p.rho0 == p.params["volume_fraction"])


HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "am_ml_suite_amd", "am_topopt")


def test_cli_setup_matches_python_and_reference_counts():
    """am_topopt's C++ BRK-A-01 set-up (problem.cpp) against the Python restatement and the
    reference's data/cpp_test_run/metadata.json counts (8151 / 784 / 65 at 4 mm)."""
    import json
    import subprocess
    G = json.load(open(os.path.join(HERE, "golden", "reference_runs.json")))["cpp_test_run"]
    for res in (4.0, 12.0, 6.0):
        out = subprocess.run([CLI, "--dump-domain", "--resolution", str(res)],
                             capture_output=True, text=True, check=True).stdout
        d = json.loads(out)
        dom = A.bracket_domain(res)
        lc = A.bracket_load_case(dom)
        assert tuple(d["domain_shape"]) == dom.shape
        assert (d["n_design"], d["n_fixed"], d["n_void"]) == (dom.n_design, dom.n_fixed,
                                                               dom.n_void)
        assert d["n_dofs"] == lc.n_dofs and d["n_load_nodes"] == len(lc.forces)
        assert d["first_load_node"] == lc.forces[0][0]
        assert d["total_fz"] == pytest.approx(-15000.0)
        assert d["n_fixed_dofs"] == len(lc.constrained)
        if res == 4.0:
            assert [d["n_design"], d["n_fixed"], d["n_void"]] == [G["n_design"], G["n_fixed"],
                                                                  G["n_void"]]
            assert d["domain_shape"] == G["domain_shape"]


def test_cli_arguments():
    import subprocess
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True)
    assert r.returncode == 1 and "--resolution <mm>" in r.stdout and "--output_dir" in r.stdout
    r = subprocess.run([CLI, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "Unknown option: --bogus" in r.stderr
    r = subprocess.run([CLI, "--resolution", "abc", "--dump-domain"], capture_output=True,
                       text=True)
    assert r.returncode == 1
