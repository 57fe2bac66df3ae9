"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/hfgpu.h declares, the host-side setup matches the oracle bit for bit,
and misuse is reported through return codes (no exception crosses the ABI)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import helfem_amd as H
import oracle as O
from helfem_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for hdr in sorted((ROOT / "include").glob("*.h")):
        txt = re.sub(r"/\*.*?\*/", "", hdr.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hf_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers them all
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version_string():
    assert b"gfx950" in _lib.lib().hf_version()


@pytest.mark.parametrize("Z,lmax,mmax,nel", [(10, 1, 1, 5), (54, 2, 1, 3)])
def test_host_setup_matches_oracle_bitwise(Z, lmax, mmax, nel):
    hb = H.AtomicBasis(Z, lmax, mmax, nel)
    ob = O.AtomicBasis(Z, lmax, mmax, nel)
    assert (hb.Nbf, hb.Nrad, hb.Nel, hb.N_L) == (ob.Nbf, ob.Nrad, ob.Nel, ob.N_L)
    assert np.array_equal(hb.first, ob.first) and np.array_equal(hb.nprim, ob.nprim)
    hb.compute_tei()
    ob.compute_tei()
    for kind in range(3):
        for L in range(hb.N_L):
            for e in range(hb.Nel):
                a, b = hb.get_cache(kind, L, e), ob.cache(kind, L, e)
                assert a.shape == b.shape and np.array_equal(a, b), (kind, L, e)
    assert np.array_equal(hb.overlap(), ob.overlap())
    assert np.array_equal(hb.kinetic(), ob.kinetic())
    assert np.array_equal(hb.nuclear(), ob.nuclear())


def test_error_codes():
    L = _lib.lib()
    h = ctypes.c_void_p()
    bval = np.array([0.0, 1.0, 0.5])          # not increasing
    lv = np.array([0], np.int32)
    rc = L.hf_atomic_create(None, 2, 15, 0, bval.ctypes.data, 3, lv.ctypes.data, lv.ctypes.data, 1, ctypes.byref(h))
    assert rc == -1 and b"increase" in L.hf_last_error()
    bad_m = np.array([2], np.int32)
    rc = L.hf_atomic_create(None, 2, 15, 0, np.array([0.0, 1.0]).ctypes.data, 2, lv.ctypes.data,
                            bad_m.ctypes.data, 1, ctypes.byref(h))
    assert rc == -1
    b = H.AtomicBasis(2, 0, 0, 2)
    with pytest.raises(H.HfError) as ei:           # before compute_tei
        b.get_cache(2, 0, 0)
    assert ei.value.code == -2
    with pytest.raises(H.HfError) as ei:           # host-only handle has no device
        b.coulomb(np.eye(b.Nbf))
    assert ei.value.code == -1
    assert b"without a device context" in L.hf_last_error()
    with pytest.raises(H.HfError):                 # out-of-range cache index
        b.set_cache(0, 5, 0, np.eye(14))
    b.set_cache(0, 0, 0, np.eye(3))                # wrong shape is caught at commit
    with pytest.raises(H.HfError):
        b.commit_tei()


def test_exact_exchange_fractions_through_the_abi():
    """hf_xc_exact_exchange is host-only: the hybrids' fractions match the
    oracle's, pure functionals give zeros, unknown ids are refused."""
    import oracle as O
    from helfem_amd.dftgrid import exact_exchange
    for xf in (0, 1, 101, 106, 202, 402, 406, 433, 470):
        assert exact_exchange(xf) == O.exact_exchange(xf)
    with pytest.raises(RuntimeError):
        exact_exchange(999)
