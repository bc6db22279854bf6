"""Full-size parity on the single-complex configurations (bit-exact against the oracle)."""
import numpy as np
import pytest

from oracle import oracle as O

import specseqplusplus_amd as S
from specseqplusplus_amd import configs as CF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    hd = S.Handle(0)
    yield hd
    hd.close()


@pytest.fixture(scope="module")
def c2():
    m = S.vr_build(CF.c2_points(), 1.0, 2)
    return m, O.Matrix(m.colptr, m.rows, m.dims)


@pytest.fixture(scope="module")
def c3():
    from specseqplusplus_amd.alpha import alpha_complex, c3_points
    m = alpha_complex(c3_points(), 90.0)
    return m, O.Matrix(m.colptr, m.rows, m.dims)


def test_c2_model3(h, c2):
    m, mo = c2
    pairs, ess, st = h.reduce(m, S.params(**CF.C2_PARAMS), stats=True)
    ref = O.twist_pairs(mo, skip=O.collapse_skip(mo))
    assert np.array_equal(pairs, ref)
    assert st["n_kept_edges"] == int((O.collapse_skip(mo)[mo.dims == 1] == 0).sum())


def test_c2_model3_dualize_equals(h, c2):
    m, mo = c2
    p3 = h.reduce(m, S.params(**CF.C2_PARAMS))[0]
    kw = dict(CF.C2_PARAMS)
    kw["dualize"] = 1
    assert np.array_equal(h.reduce(m, S.params(**kw))[0], p3)


def test_c2_full_complex_dualize(h, c2):
    m, mo = c2
    pairs = h.reduce(m, S.params(dualize=1))[0]
    assert np.array_equal(pairs, O.twist_pairs(mo))


def test_c3_alpha(h, c3):
    m, mo = c3
    ref = O.twist_pairs(mo)
    for kw in [dict(model=0, queue_size=20000, level1=1000, level2=16000), dict(dualize=1)]:
        pairs = h.reduce(m, S.params(**kw))[0]
        assert np.array_equal(pairs, ref), kw
