"""CPU tests of the shipped library's host side: the native symbolic analysis equals the oracle's (bit-exact integer
structure), the C ABI loads and exports every entry point of include/ampsolve.h, argument and state errors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import ampsolve as amp
from ampsolve import _capi
from ampsolve import workloads as W
from fixtures import chain, grid27, hex_matrix, ldlt_system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ["perm", "pinv", "parent", "colCount", "lowerP", "lowerI", "lowerSrc", "superStart", "colSuper",
          "superParent", "childHead", "childNext", "Sp", "Si", "Lxp"]

PATTERNS = {
    "poisson2d_64": lambda: W.poisson2d(64)[0],
    "heat3d_12": lambda: W.heat3d(12)[0],
    "grid27_9x7x5": lambda: grid27(9, 7, 5)[0],
    "hex": lambda: hex_matrix()[0],
    "chain": lambda: chain(97)[0],
    "ldlt_system": lambda: ldlt_system(30)[0],
}


@pytest.mark.parametrize("name", list(PATTERNS))
@pytest.mark.parametrize("relax", [True, False])
def test_native_analyze_equals_oracle(name, relax):
    pat = PATTERNS[name]()
    s = amp.analyze(pat, {"relax": relax})
    o = O.Sym(pat, relax=relax)
    for f in FIELDS:
        assert np.array_equal(getattr(s, f), getattr(o, f)), f
    assert (s.nnzL, s.nnzA, s.nnzLStored, s.maxFront, s.flops) == (o.nnzL, o.nnzA, o.nnzLStored, o.maxFront, o.flops)
    assert s.patternKey == o.patternKey == amp.patternFingerprint(pat)


def test_native_amd_and_given_perm():
    pat = W.heat3d(8)[0]
    assert np.array_equal(amp.amdOrder(pat), O.amd(pat))
    rng = np.random.default_rng(3)
    perm = rng.permutation(pat["n"]).astype(np.int32)
    s = amp.analyze(pat, {"perm": perm})
    o = O.Sym(pat, perm=perm)
    for f in FIELDS:
        assert np.array_equal(getattr(s, f), getattr(o, f)), f


def test_every_declared_entry_point_is_exported():
    header = open(os.path.join(ROOT, "include", "ampsolve.h")).read()
    declared = set(re.findall(r"\b(amp_[a-z_0-9]+)\s*\(", header))
    assert declared, "no declarations found"
    lib = C.CDLL(_capi.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared but not exported"
    assert declared == set(_capi.SIGNATURES), "binding and header disagree"


def test_argument_errors_without_device():
    L = _capi.lib()
    out = C.c_void_p()
    assert L.amp_analyze(-1, np.zeros(1, np.int32), np.zeros(1, np.int32), None, 1, C.byref(out)) == _capi.AMP_ERR_ARG
    assert "bad arguments" in L.amp_last_error().decode()
    with pytest.raises(ValueError, match="unknown real factor"):
        amp.setDefaultRealFactor("fast")
    pat = grid27(5, 5, 5)[0]
    sym = amp.analyze(pat)
    with pytest.raises(ValueError, match="values for a pattern"):
        amp.factorNumeric(sym, np.zeros(pat["nnz"] - 1))
    with pytest.raises(ValueError, match="perm has length"):
        amp.analyze(pat, {"perm": np.arange(3)})


def test_plan_rejects_fronts_beyond_tile_offsets():
    """The product kernels address a tile's operands with 32-bit offsets, so a super-panel so wide that a tiled
    front's rows x width reaches 2^31 is refused with an argument error (before any device work)."""
    pat = W.heat3d(12)[0]
    sym = amp.analyze(pat)
    with pytest.raises(Exception, match="too tall for 32-bit tile offsets"):
        sym.set_params({"smallMax": 96, "smallNarrow": 32, "wideLaunch": 512, "superPanel": 1 << 24})


def test_sym_import_roundtrip():
    """The analysis handed over by a host symbolic phase gives the same structure as the native one."""
    pat = grid27(8, 7, 6)[0]
    s = amp.analyze(pat)
    d = _capi.SymDesc()
    d.n, d.nsuper, d.nnzPattern = s.n, s.nsuper, pat["nnz"]
    keep = {}
    for f in ("perm", "parent", "colCount", "superStart", "Sp", "Si", "lowerP", "lowerI", "lowerSrc"):
        keep[f] = np.ascontiguousarray(getattr(s, f), dtype=np.int32)
        setattr(d, f, keep[f].ctypes.data)
    d.patternKey = s.patternKey.encode()
    h = C.c_void_p()
    _capi.check(_capi.lib().amp_sym_import(C.byref(d), C.byref(h)))
    t = amp.Symbolic(h.value)
    for f in FIELDS:
        assert np.array_equal(getattr(t, f), getattr(s, f)), f
    assert t.flops == s.flops and t.gemmFlops == s.gemmFlops


def _import_desc(s, nnz, overrides):
    d = _capi.SymDesc()
    d.n, d.nsuper, d.nnzPattern = s.n, overrides.pop("nsuper", s.nsuper), overrides.pop("nnzPattern", nnz)
    keep = {}
    for f in ("perm", "parent", "colCount", "superStart", "Sp", "Si", "lowerP", "lowerI", "lowerSrc"):
        keep[f] = np.ascontiguousarray(overrides.get(f, getattr(s, f)), dtype=np.int32)
        setattr(d, f, keep[f].ctypes.data)
    return d, keep


def _corrupt(s, field, fn):
    a = np.array(getattr(s, field), dtype=np.int32, copy=True)
    fn(a)
    return {field: a}


CORRUPTIONS = {
    "perm duplicate": lambda s: _corrupt(s, "perm", lambda a: a.__setitem__(1, a[0])),
    "parent below child": lambda s: _corrupt(s, "parent", lambda a: a.__setitem__(5, 2)),
    "parent out of range": lambda s: _corrupt(s, "parent", lambda a: a.__setitem__(0, s.n + 7)),
    "superStart not increasing": lambda s: _corrupt(s, "superStart", lambda a: a.__setitem__(2, a[1])),
    "Sp not increasing": lambda s: _corrupt(s, "Sp", lambda a: a.__setitem__(3, a[2] - 1)),
    "Si row out of range": lambda s: _corrupt(s, "Si", lambda a: a.__setitem__(s.Sp[1] - 1, s.n + 3)),
    "Si own column": lambda s: _corrupt(s, "Si", lambda a: a.__setitem__(0, 1)),
    "lowerI out of range": lambda s: _corrupt(s, "lowerI", lambda a: a.__setitem__(len(a) - 1, s.n)),
    "lowerI off the structure": lambda s: _corrupt(s, "lowerI", lambda a: a.__setitem__(0, s.n - 1)),
    "lowerSrc out of range": lambda s: _corrupt(s, "lowerSrc", lambda a: a.__setitem__(3, 1 << 30)),
    "lowerP not increasing": lambda s: _corrupt(s, "lowerP", lambda a: a.__setitem__(4, a[3] - 1)),
    "colCount zero": lambda s: _corrupt(s, "colCount", lambda a: a.__setitem__(0, 0)),
    "no supernodes": lambda s: {"nsuper": 0},
}


@pytest.mark.parametrize("what", list(CORRUPTIONS))
def test_sym_import_rejects_corrupted_descriptors(what):
    """Every index the plans and the kernels follow is range-checked at the boundary: a corrupted analysis is an
    argument error, never an out-of-range access."""
    pat = grid27(6, 5, 4)[0]
    s = amp.analyze(pat)
    d, keep = _import_desc(s, pat["nnz"], CORRUPTIONS[what](s))
    h = C.c_void_p()
    rc = _capi.lib().amp_sym_import(C.byref(d), C.byref(h))
    assert rc == _capi.AMP_ERR_ARG, (what, rc)
    assert not h.value
    assert _capi.lib().amp_last_error().decode().startswith("amp_sym_import")


def test_analyze_nd_ordering():
    """ordering 'nd' (sparse.mjs:678-679): needs node coordinates of length n (the reference's message), takes the
    dissection order as its permutation, and reports the ordering time apart from the symbolic time."""
    import ampsolve as amp
    from test_oracle import grid_pattern

    pat, x, y = grid_pattern(30, 20)
    with pytest.raises(ValueError, match="needs node coordinates"):
        amp.analyze(pat, {"ordering": "nd"})
    with pytest.raises(ValueError, match="needs node coordinates"):
        amp.analyze(pat, {"ordering": "nd", "x": x[:-1], "y": y})
    sym = amp.analyze(pat, {"ordering": "nd", "x": x, "y": y})
    given = amp.analyze(pat, {"perm": amp.ndOrder(pat, x, y)})  # the analysis postorders the given order
    assert np.array_equal(sym.perm, given.perm) and sym.nnzL == given.nnzL
    assert set(sym.timing) == {"ordering_ms", "symbolic_ms"} and sym.timing["ordering_ms"] >= 0
    assert given.timing["ordering_ms"] == 0
    So = O.Sym(pat, perm=O.nd_order(pat, x, y))
    assert np.array_equal(sym.perm, So.perm) and np.array_equal(sym.Si, So.Si)
    assert np.array_equal(amp.analyze(pat).perm, amp.analyze(pat, {"perm": amp.amdOrder(pat)}).perm)
    ch = amp.SparseCholesky(pat, ordering="nd", x=x, y=y)
    assert np.array_equal(ch.perm, sym.perm) and "ordering_ms" in ch.timing
    with pytest.raises(Exception):
        amp.ndOrder(pat, x, y, leaf=0)
