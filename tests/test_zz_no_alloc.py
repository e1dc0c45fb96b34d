"""alpsg_disp_batch_noalloc (no device allocation: buffers carved from the
library's fixed per-GPU scratch array) against alpsg_disp_batch on the same
handle, through the C-ABI: two calls in a row, the second smaller, det D to
1e-11 of the determinant's term scale; batches beyond the scratch and
handles it does not take are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _call(E, om):
    from alps_amd.engine import _dp, load_library
    om = np.ascontiguousarray(np.asarray(om, dtype=np.complex128).ravel())
    det = np.zeros(om.size, dtype=np.complex128)
    rc = load_library().alpsg_disp_batch_noalloc(E._h, om.size, _dp(om), _dp(det))
    return rc, det


def _grid(n):
    """n omegas over resonant and non-resonant rows, both signs of Im."""
    t = (np.arange(n) + 0.5) / n
    return 0.05 + 1.9 * t + 1j * (-0.3 + 0.32 * np.mod(7.0 * t, 1.0))


@pytest.mark.parametrize("case,k", [("small_bimax", (0.3, 0.3)), ("three_species", (0.5, 0.2)),
                                    ("config1", (0.5, 0.5))])
def test_noalloc_matches_disp_batch_on_two_calls(case, k):
    from alps_amd.engine import Engine
    p = getattr(cases, case)()
    E = Engine(p)
    E.set_k(*k)
    print("case", case, flush=True)
    for om in (_grid(1500), _grid(7)):   # the second call is smaller (and takes the small-batch sweep)
        rc, det = _call(E, om)
        assert rc == 0
        ref, _, _, wave = E.disp_batch(om, want_tensors=True)
        for i in range(om.size):
            err = abs(det[i] - ref[i]) / cases.det_scale(wave[i])
            assert err <= TOL, (case, om.size, i, err)
    E.close()


def test_noalloc_refuses_what_it_cannot_take():
    from alps_amd.engine import Engine, load_library
    L = load_library()
    p = cases.small_bimax()
    E = Engine(p)
    E.set_k(0.3, 0.3)
    rc, _ = _call(E, _grid(400000))   # far beyond the 32 MiB of scratch
    assert rc == 1 and b"scratch" in L.alpsg_last_error()
    # the handle still works afterwards, on both paths
    rc, det = _call(E, _grid(3))
    assert rc == 0 and np.allclose(det, E.disp_batch(_grid(3)), rtol=1e-12, atol=0)
    E.close()
    R = Engine(cases.rel_pair())
    R.set_k(0.3, 0.1)
    rc, _ = _call(R, [0.5 + 0.05j])
    assert rc == 2   # ALPSG_E_UNSUPPORTED: relativistic species take the handle's buffers
    R.close()
