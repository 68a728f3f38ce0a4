"""The JavaScript host layer (applied-multiphysics_amd/js/ampsolve.mjs over the Node-API addon): same analysis as the
oracle, the reference's error messages, and (GPU) factor / solve / LDL^T / refinement / a handed-over analysis."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from ampsolve import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "applied-multiphysics_amd", "js", "test", "run.mjs")
ADDON = os.path.join(ROOT, "applied-multiphysics_amd", "js", "amp_native.node")
needs_node = pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")


def laplace(k):
    I = sp.identity(k)
    T = sp.diags([-np.ones(k - 1), 2 * np.ones(k), -np.ones(k - 1)], [-1, 0, 1])
    A = (sp.kron(I, T) + sp.kron(T, I)).tocsr()
    A.sort_indices()
    return W.csr_dict(A), A.data


def run(mode):
    assert os.path.exists(ADDON), "amp_native.node not built"
    p = subprocess.run(["node", RUN, mode], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


@needs_node
def test_js_analysis_and_errors_cpu():
    out = run("--cpu")
    pat, _ = laplace(30)
    S = O.Sym(pat)
    assert out["n"] == S.n and out["nsuper"] == S.nsuper
    assert out["perm"] == S.perm.tolist() and out["Sp"] == S.Sp.tolist()
    assert out["Lxp_last"] == S.nnzLStored
    assert out["patternKey"] == out["fingerprint"] == S.patternKey
    assert out["errors"][0].startswith("unknown real factor fast")
    assert "values for a pattern with" in out["errors"][1]
    assert out["errors"][2] == "symbolic analysis is for a different sparsity pattern"
    # wrong element type, short colidx, rowptr shorter than n + 1: refused by the addon before the library reads
    assert len(out["addon_errors"]) == 3
    assert "shorter than the pattern" in out["addon_errors"][1]
    assert "shorter than the pattern" in out["addon_errors"][2]
    # orderings: the dissection order equals the oracle's, analyze('nd') postorders it, AMD equals the oracle's
    g = np.arange(pat["n"])
    x, y = (g % 30).astype(float), (g // 30).astype(float)
    nd = O.nd_order(pat, x, y)
    assert out["nd"] == nd.tolist() and out["amd"] == O.amd(pat).tolist()
    assert out["nd_perm"] == O.Sym(pat, perm=nd).perm.tolist()
    assert out["nd_timing"] == ["ordering_ms", "symbolic_ms"]
    assert out["nd_error"] == "analyze: ordering nd needs node coordinates x and y of length n"
    assert out["kinds"] == ["hip", "hip-f32", "hip-bf16x3"]


@needs_node
@pytest.mark.gpu
def test_js_factor_solve_gpu():
    out = run("--gpu")
    assert out["ok"] is True and out["relres"] < 1e-12
    assert out["multi"] < 1e-12
    assert out["short_b"] == "solve: b / out shorter than n x nrhs"
    assert out["indefinite_ok"] is False
    assert out["solve_error"] == "SparseCholesky: solve without a successful factorization"
    assert out["ldlt"]["ok"] and out["ldlt"]["nNeg"] > 0
    assert out["ldlt_relres"] < 1e-10
    assert out["imported_relres"] < 1e-12
    assert out["f32_refined"] < 1e-12
