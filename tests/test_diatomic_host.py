"""Host-side diatomic setup of libhfgpu against the oracle (no GPU).

The product's compute_tei, maps, pure indices and one-electron matrices
(src/diatomic/basis.cpp:525-647, 717-737, 1032-1156, 1382-1547) must equal
the oracle's bit for bit, so that the GPU parity tests isolate the kernels.
"""
import numpy as np
import pytest

import helfem_amd as H
import oracle as O

CASES = {
    "H2-R1.4-l21-3el": (1, 1, 1.4, [2, 1], 3),
    "N2-R2.074-l432-3el": (7, 7, 2.074, [4, 3, 2], 3),
    "LiH-R3-l32-4el": (3, 1, 3.0, [3, 2], 4),
    "HeH+-R1.5-l0-1el": (2, 1, 1.5, [0], 1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_setup_bitwise_equal(name):
    Z1, Z2, R, lmm, nel = CASES[name]
    ob = O.DiatomicBasis(Z1, Z2, R, lmm, nel)
    hb = H.DiatomicBasis(Z1, Z2, R, lmm, nel)
    assert np.array_equal(ob.bval, hb.bval)
    assert (ob.Nbf, ob.Nrad, ob.Nang, ob.Nel) == (hb.Nbf, hb.Nrad, hb.Nang, hb.Nel)
    assert np.array_equal(ob.pure_idx, hb.pure_idx)
    lo, Lo = ob.maps()
    lh, Lh = hb.maps()
    assert np.array_equal(lo, lh) and np.array_equal(Lo, Lh)
    for a, b in ((ob.overlap(), hb.overlap()), (ob.kinetic(), hb.kinetic()), (ob.nuclear(), hb.nuclear())):
        assert np.array_equal(a, b)
    ob.compute_tei()
    hb.compute_tei()
    for ilm in range(ob.n_lm):
        for e in range(ob.Nel):
            for kind in range(6):
                a, b = ob.cache(kind, ilm, e), hb.get_cache(kind, ilm, e)
                assert a.shape == b.shape and np.array_equal(a, b), (kind, ilm, e)


def test_element_grid_matches_oracle():
    for igrid, zexp in ((1, 1.0), (2, 1.0), (3, 1.5), (4, 1.0), (4, 2.0)):
        assert np.array_equal(H.element_grid(4.3455, 8, igrid, zexp), O.grid(4.3455, 8, igrid, zexp))


def test_cfg_sizes_n2():
    """N2 at lmax 16, m <= 3, 8 elements: Nrad 112, Nang 107, Nbf 11894."""
    lv, mv = H.diatomic_shells([16, 16, 16, 16])
    hb = H.DiatomicBasis(7, 7, 2.074, lval=lv, mval=mv, nelem=8)
    assert (hb.Nrad, hb.Nang, hb.Ndummy, hb.Nbf) == (112, 107, 11984, 11894)
    assert abs(hb.bval[-1] - 4.3455) < 1e-4


def test_errors_are_reported():
    with pytest.raises(H.HfError):
        H.DiatomicBasis(1, 1, -1.0, [1], 2)
    hb = H.DiatomicBasis(1, 1, 1.4, [1], 2)
    with pytest.raises(H.HfError):
        hb.get_cache(0, 0, 0)  # caches not built yet
