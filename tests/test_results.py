"""Results / benchmark wire formats (src/algorithms/dmrg_results_json.mojo,
src/tests/algorithms/dmrg/dmrg_json_loader.mojo,
src/tests/benchmarks/bench_contractions.mojo:221-253)."""
import json

import numpy as np
import pytest
import torch

import chemtensor_mojo_amd as ct


def _params(tol):
    return ct.DMRGJsonParams(nsites=16, d=2, J=1.0, D=1.0, h=0.0, num_sweeps=4, maxiter_lanczos=20, chi_max=32,
                             tol_split=tol)


def test_results_json_layout_two_site():
    """Key order and layout of dmrg_results_to_json_string (dmrg_results_json.mojo:45-79)."""
    s = ct.dmrg_results_to_json_string("hip", "heisenberg_xxz", _params(1e-10), -15.0, [-14.5, -15.0],
                                       [0.5, 0.25], [1, 2, 1], 1.0)
    expected = ('{\n'
                '  "impl": "hip",\n'
                '  "model": "heisenberg_xxz",\n'
                '  "params": {\n'
                '    "nsites": 16,\n'
                '    "d": 2,\n'
                '    "J": 1.0,\n'
                '    "D": 1.0,\n'
                '    "h": 0.0,\n'
                '    "num_sweeps": 4,\n'
                '    "maxiter_lanczos": 20,\n'
                '    "chi_max": 32,\n'
                '    "tol_split": 1e-10\n'
                '  },\n'
                '  "energy_final": -15.0,\n'
                '  "en_sweeps": [-14.5, -15.0],\n'
                '  "entropy": [0.5, 0.25],\n'
                '  "bond_dims": [1, 2, 1],\n'
                '  "norm": 1.0\n'
                '}\n')
    assert s == expected
    assert json.loads(s)["params"]["tol_split"] == 1e-10


def test_results_json_optional_keys_omitted():
    """tol_split < 0 (single-site runs) and an empty entropy list drop their keys."""
    d = json.loads(ct.dmrg_results_to_json_string("hip", "m", _params(-1.0), -1.0, [-1.0], [], [1, 1], 1.0))
    assert "tol_split" not in d["params"] and "entropy" not in d
    assert list(d.keys()) == ["impl", "model", "params", "energy_final", "en_sweeps", "bond_dims", "norm"]


def test_results_json_round_trip(tmp_path):
    p = tmp_path / "r.json"
    e = [-14.984307861, -14.999999999999996]
    ct.save_dmrg_results_to_json(str(p), "hip", "heisenberg_xxz", _params(1e-10), e[-1], e, [0.1], [1, 2, 4, 2, 1],
                                 0.9999999999999999)
    r = ct.load_dmrg_reference(str(p))
    assert r.impl == "hip" and r.nsites == 16 and r.chi_max == 32 and r.tol_split == 1e-10
    assert r.en_sweeps == e and r.energy_final == e[-1]   # shortest round-trip floats: exact
    assert r.bond_dims == [1, 2, 4, 2, 1] and r.entropy == [0.1] and r.norm == 0.9999999999999999
    p2 = tmp_path / "s.json"
    ct.save_dmrg_results_to_json(str(p2), "hip", "m", _params(-1.0), -1.0, [-1.0], [], [1, 1], 1.0)
    r2 = ct.load_dmrg_reference(str(p2))
    assert r2.tol_split == -1.0 and r2.entropy == []


def test_bench_record_jsonl(tmp_path):
    p = tmp_path / "sub" / "bench.jsonl"
    ct.append_bench_record(str(p), "heff_apply", 16, 2, 1024, 0.00057, -1.5, 20)
    ct.append_bench_record(str(p), "dmrg_sweep", 24, 2, 1024, 0.70, -36.0, 1)
    lines = p.read_text().splitlines()
    assert len(lines) == 2
    assert lines[0] == ('{"backend":"hip","operation":"heff_apply","params":{"nsites":16,"d":2,"chi_max":1024},'
                        '"time_seconds":0.00057,"result":-1.5,"runs":20}')
    assert json.loads(lines[1])["params"]["nsites"] == 24


def _mps_from_state(psi, L, d=2):
    """Exact MPS of a state vector by successive numpy SVDs (independent of the library)."""
    sites, rest, Dl = [], psi.reshape(1, -1), 1
    for i in range(L - 1):
        M = rest.reshape(Dl * d, -1)
        U, S, Vt = np.linalg.svd(M, full_matrices=False)
        k = int(np.sum(S > 1e-14 * S[0]))
        sites.append(U[:, :k].reshape(Dl, d, k))
        rest = S[:k, None] * Vt[:k]
        Dl = k
    sites.append(rest.reshape(Dl, d, 1))
    return sites


@pytest.mark.gpu
def test_bond_entropies_and_norm_vs_state_vector(ctx):
    """Entropies and norm of a random 8-site state (scaled by 3) against numpy
    SVDs of the state vector at every cut; a product state gives zeros."""
    L = 8
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(2 ** L) * 3.0
    sites = _mps_from_state(psi, L)
    # give the MPS a non-canonical gauge: scale one site up and its neighbour down
    sites[3] = sites[3] * 2.0
    sites[4] = sites[4] * 0.5
    mps = ct.mps_from_tensors(ctx, sites)
    assert abs(ct.mps_norm(mps, ctx) - np.linalg.norm(psi)) < 1e-12 * np.linalg.norm(psi)
    ent = ct.bond_entropies(mps, ctx)
    exp = []
    for cut in range(1, L):
        s = np.linalg.svd(psi.reshape(2 ** cut, -1), compute_uv=False)
        p = s ** 2 / np.sum(s ** 2)
        p = p[p > 0]
        exp.append(float(-(p * np.log(p)).sum()))
    assert np.allclose(ent, exp, atol=1e-10)
    prod = ct.create_product_mps(ctx, 2, [0, 1, 1, 0])
    assert np.allclose(ct.bond_entropies(prod, ctx), 0.0, atol=1e-12)
