"""The allocation-free element dump (tf_element_dump_range_noalloc) against tf_element_dump, through
the C ABI: two calls in a row on the same GPU array, the second with fewer prisms, each equal
bitwise to one tf_element_dump call per prism; more prisms than the static scratch holds is
rejected with TF_INVALID."""
import numpy as np
import pytest

from tissuefolding_paperversion_amd import _lib as L
from tissuefolding_paperversion_amd import configs

pytestmark = pytest.mark.gpu


def _dump_range(s, e0, n, dt):
    fe, fv = np.full(18 * n, np.nan), np.full(18 * n, np.nan)
    Ke, Kv = np.full(324 * n, np.nan), np.full(324 * n, np.nan)
    rc = s._lib.tf_element_dump_range_noalloc(s.h, e0, n, dt, L.ptr(fe), L.ptr(fv), L.ptr(Ke), L.ptr(Kv))
    return rc, fe.reshape(n, 18), fv.reshape(n, 18), Ke.reshape(n, 18, 18), Kv.reshape(n, 18, 18)


def test_noalloc_element_dump_matches_the_allocating_one():
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    t = configs.wing_disc_tissue(a=40.0, b=25.0, h=3.66, z_levels=(0.0, 2.0, 7.0, 12.0), seed=2, growth_seed=3,
                                 name="dump ellipse")
    t.eta_int = np.full(t.n_elements, 40.0)  # a viscous tangent that is not zero
    t.rates = t.rates * 10.0
    s = NewtonRaphsonSolver(t)
    s.runOneStep(t.dt)  # uk != un: the viscous forces are not zero either
    for e0, n in ((3, 300), (250, 17)):  # the second call has fewer prisms
        print(f"noalloc dump: prisms {e0}..{e0 + n - 1}", flush=True)
        rc, fe, fv, Ke, Kv = _dump_range(s, e0, n, t.dt)
        assert rc == L.TF_OK, s._lib.tf_last_error(s.h).decode()
        for k in range(n):
            ref = s.element_dump(e0 + k, t.dt)
            assert np.array_equal(fe[k], ref["fe"]) and np.array_equal(fv[k], ref["fv"])
            assert np.array_equal(Ke[k], ref["Ke"]) and np.array_equal(Kv[k], ref["Kv"])
        assert np.abs(Kv).max() > 0 and np.abs(fv).max() > 0
    # null outputs are allowed
    rc = s._lib.tf_element_dump_range_noalloc(s.h, 0, 5, t.dt, None, None, None, None)
    assert rc == L.TF_OK
    # beyond the scratch, and outside the mesh
    assert s._lib.tf_element_dump_range_noalloc(s.h, 0, 513, t.dt, None, None, None, None) == L.TF_INVALID
    assert "at most 512 prisms" in s._lib.tf_last_error(s.h).decode()
    assert s._lib.tf_element_dump_range_noalloc(s.h, t.n_elements - 3, 4, t.dt, None, None, None, None) == L.TF_INVALID
    assert s._lib.tf_element_dump_range_noalloc(s.h, 0, 0, t.dt, None, None, None, None) == L.TF_INVALID
