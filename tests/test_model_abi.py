"""CPU tests of the product's host side: the C ABI exports and the host
problem model against the oracle (bitwise: same constants, same operation
order, same channel order)."""
import os
import re

import numpy as np
import pytest

from coupled_cluster_doubles_amd import CCDError, Model
from coupled_cluster_doubles_amd import _lib
from oracle.pyoracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "ccd_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ccd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = _declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # and the ctypes table covers every declaration
    assert sorted(s[0] for s in _lib.SIGNATURES) == names


CONFIGS = [
    ("pairing", (8, 8, 0.5, 1.0)),
    ("pairing", (16, 16, 0.3, 1.0)),
    ("heg", (1.0, 3, 2)),
    ("heg", (2.0, 6, 3)),
    ("heg", (1.0, 19, 2)),
    ("infmatter", (0.08, 4, 2)),
    ("infmatter", (0.16, 3, 2)),
]


def _both(kind, args, load=False):
    if kind == "pairing":
        return Model.pairing(*args), Oracle.pairing(*args, load=load)
    if kind == "infmatter":
        return Model.inf_matter(*args), Oracle.inf_matter(*args, load=load)
    return Model.electron_gas(*args), Oracle.electron_gas(*args, load=load)


@pytest.mark.parametrize("kind,args", CONFIGS)
def test_model_matches_oracle_bitwise(kind, args):
    m, o = _both(kind, args)
    assert m.n_sp == o.nsp and m.n_holes == o.nh and m.n_t2 == o.nt2
    assert m.eref() == o.eref()
    assert np.array_equal(m.eps(), o.eps())
    d, od = m.dims(), o.dims()
    for k, ok in (("pp_dim", "pp"), ("hh_dim", "hh"), ("ph_dim", "mph"), ("hp_dim", "mhp")):
        assert np.array_equal(d[k], od[ok]), k
    assert np.array_equal(m.t2_labels(), o.t2_labels())


@pytest.mark.parametrize("kind,args", CONFIGS[:4] + CONFIGS[5:6])
def test_model_host_v_matches_oracle(kind, args):
    m, o = _both(kind, args, load=True)
    p = m.problem(host_v=True).contents
    import ctypes as C
    v = np.ctypeslib.as_array(C.cast(p.v_hhpp, C.POINTER(C.c_double)), (m.n_t2,))
    assert np.array_equal(v, o.vhhpp())


def test_model_rejects_bad_arguments():
    with pytest.raises(CCDError):
        Model.pairing(0, 8, 0.5, 1.0)
    with pytest.raises(CCDError):
        Model.electron_gas(-1.0, 4, 2)
    with pytest.raises(CCDError):
        Model.electron_gas(1.0, 2, 3)


def test_headline_sizes():
    m = Model.electron_gas(1.0, 32, 4)
    assert (m.n_sp, m.n_holes, m.n_t2) == (1850, 54, 2925948)


def test_context_rejects_bad_modes_before_touching_a_device():
    """Argument errors come back as CCD_E_INVALID from host validation."""
    from coupled_cluster_doubles_amd import FP64, MATRIX_FREE, Context, _lib
    m = Model.electron_gas(1.0, 4, 2)
    with pytest.raises(CCDError, match="generator"):
        Context(m, FP64, host_v=True, matrix_free=True)   # nothing to regenerate V from
    with pytest.raises(CCDError, match="precision"):
        Context(m, 7)
    h = _lib.C.c_void_p()
    assert _lib.lib().ccd_ctx_create(m.problem(True), FP64 | MATRIX_FREE, _lib.C.byref(h)) == -1


def test_inf_matter_model_rejects_symmetric_matter():
    with pytest.raises(CCDError, match="pure neutron matter"):
        Model.inf_matter(0.16, 4, 2, tz_max=2)



def test_batch_rejects_mixed_interaction_sources_before_touching_a_device():
    import ctypes as C
    from coupled_cluster_doubles_amd import Batch
    from coupled_cluster_doubles_amd._lib import check, lib
    a, b = Model.pairing(8, 8, 0.5, 1.0), Model.inf_matter(0.08, 3, 2)
    probs = (C.c_void_p * 2)(C.cast(a.problem(), C.c_void_p), C.cast(b.problem(host_v=True), C.c_void_p))
    h = C.c_void_p()
    with pytest.raises(CCDError, match="generated V for every problem or for none"):
        check(lib().ccd_batch_create(probs, 2, 0, C.byref(h)), "ccd_batch_create")
    with pytest.raises(CCDError, match="no problems"):
        Batch([])


def test_sweep_is_seeded_and_covers_the_bases():
    from coupled_cluster_doubles_amd.sweep import sweep_configs
    a, b = sweep_configs(), sweep_configs()
    assert a == b and len(a) == 64
    heg = [x for k, x in a if k == "heg"]
    pair = [x for k, x in a if k == "pairing"]
    assert len(heg) == len(pair) == 32
    assert {x[1:] for x in heg} == {(19, 2), (25, 3), (32, 4)}
    assert all(0.5 <= x[0] <= 5.0 for x in heg)
    assert all(x[0] == x[1] and x[0] % 2 == 0 and 8 <= x[0] <= 16 and 0.2 <= x[2] <= 0.8 for x in pair)
