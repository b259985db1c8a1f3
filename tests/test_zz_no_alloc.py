"""crt_radiation_pressure_noalloc: the allocation-free variant of
crt_radiation_pressure (scratch in a static device array of the library),
called through the C ABI.  Its device output must equal the existing entry
point's result bit for bit on consecutive calls, the second with a smaller
grid (the scratch keeps the previous call's contents), and grids beyond the
scratch maximum must be rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from crt import capi, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def body():
    m = scenes.displaced_icosphere(6)
    return capi.Scene(m.vertices, m.indices)


def _opts(sun, grid, ref):
    o = capi.FluxOpts()
    o.sun_dir[:] = [float(x) for x in sun]
    o.pressure, o.rho_s, o.rho_d, o.grid = 4.56e-6, 0.3, 0.6, int(grid)
    o.center[:] = [0.0, 0.0, 0.0]
    o.radius = 0.0
    o.ref_point[:] = [float(x) for x in ref]
    return o


def test_noalloc_matches_radiation_pressure(body):
    L = capi.load()
    out = capi.DeviceBuffer(7 * 8)
    sun = np.array([0.3, -0.5, 0.81])
    sun /= np.linalg.norm(sun)
    ref = (120.0, -40.0, 15.0)
    for grid in (2048, 512, 1000):   # second call smaller; third not a multiple of the 8 x 8 tile
        out.fill(0xFF)
        o = _opts(sun, grid, ref)
        print("noalloc grid %d" % grid, flush=True)
        assert L.crt_radiation_pressure_noalloc(body.h, C.byref(o), C.c_void_p(out.ptr), None) == capi.CRT_OK
        capi.device_synchronize()
        got = out.to_host(np.zeros(7))
        F, T, nh, _ = body.radiation_pressure(sun, 4.56e-6, 0.3, 0.6, grid, (0, 0, 0), 0.0, ref)
        assert np.array_equal(got[:3], F) and np.array_equal(got[3:6], T), (got, F, T)
        assert got[6] == nh and nh > 0


def test_noalloc_rejects_grid_beyond_scratch(body):
    L = capi.load()
    out = capi.DeviceBuffer(7 * 8)
    o = _opts((0.0, 0.0, 1.0), 4096, (0, 0, 0))
    assert L.crt_radiation_pressure_noalloc(body.h, C.byref(o), C.c_void_p(out.ptr), None) == capi.CRT_ERR_INVALID
    assert b"2048" in L.crt_last_error()
    o = _opts((0.0, 0.0, 1.0), 0, (0, 0, 0))
    assert L.crt_radiation_pressure_noalloc(body.h, C.byref(o), C.c_void_p(out.ptr), None) == capi.CRT_ERR_INVALID
