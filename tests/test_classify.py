"""Host classification of the product (clowns_classify: set_nflag, row
numbering, mp; C++ in the solver library) against the C oracle and the
flang build of the reference, bit for bit, on small instances of every
configuration and on configuration 1 at full size (64x64x32)."""
import numpy as np
import pytest

from clowns3d import classify, configs
from test_oracle_assembly import GRIDS, build_oracle, build_ref


def _check(g, o):
    c = classify(g)
    N = o["N"]
    assert c["N"] == N
    assert np.array_equal(c["nf"], o["nf"])
    assert np.array_equal(c["nb"], o["nb"])
    assert np.array_equal(c["row2cell"], o["row2cell"][:N])
    assert np.array_equal(c["mno"], o["mno"])
    return c


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_classify_matches_oracle(name):
    g = GRIDS[name]()
    _check(g, build_oracle(g))


def test_classify_config1_matches_reference():
    g = configs.config1()
    r = build_ref(g)
    o = build_oracle(g)
    c = _check(g, o)
    assert np.array_equal(c["nf"], r["nf"]) and np.array_equal(c["row2cell"], r["row2cell"][: r["N"]])
    assert c["mp"] == r["mp"] == 0.25
    # N of configuration 1 (64x64x32, surface at 0.6H, 16x16x12 obstacle)
    assert 70000 < c["N"] < 85000


def test_classify_rejects_open_ring():
    from clowns3d import ClownsError
    g = configs.config_wavy_box(10, 8, 6, (2, 4, 2, 4, 0, 2))
    g.solid[3, 3, 0] = 0  # a hole in the wall
    with pytest.raises(ClownsError):
        classify(g)
