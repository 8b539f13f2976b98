"""GPU: the allocation-free spectral radius (hc_spectral_radius_noalloc) against the allocating
hc_spectral_radius through the C-ABI, on two calls in a row on one GPU, the second on a smaller
case (the scratch keeps the first call's partials).  The device arrays are taken with the HIP
runtime the library is bound to (looked up by its soname, so the one already loaded)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    import hyperboliconvergence_amd as hca
    hca.library()
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


def _ctx(cfg):
    import hyperboliconvergence_amd as hca
    ctx = hca.Context(0)
    ctx.set_mesh(cfg["nodes"])
    ctx.set_case(cfg["case"])
    ctx.set_boundary_refs(cfg["bref"])
    ctx.set_state(cfg["q"])
    return ctx


def test_spectral_radius_noalloc_matches_the_allocating_entry_point():
    import hyperboliconvergence_amd as hca
    from hyperboliconvergence_amd import cases
    hip = _hip()
    H2D, D2H = 1, 2
    big = cases.config_isentropic_vortex(512, hca.FLUX_AUSMPUP)   # 262144 cells: 1024 partials
    small = cases.config_cylinder(64, 32, 0.5, 10.0, flux=hca.FLUX_ROE)   # 2048 cells: 8 partials
    for cfg in (big, small):
        with _ctx(cfg) as ctx:
            res = np.ascontiguousarray(ctx.residual())
            ref = ctx.spectral_radius(res)
            d_res, d_out = C.c_void_p(), C.c_void_p()
            assert hip.hipMalloc(C.byref(d_res), res.nbytes) == 0
            assert hip.hipMalloc(C.byref(d_out), 8) == 0
            try:
                out = np.array([-1.0])
                assert hip.hipMemcpy(d_res, res.ctypes.data_as(C.c_void_p), res.nbytes, H2D) == 0
                assert hip.hipMemcpy(d_out, out.ctypes.data_as(C.c_void_p), 8, H2D) == 0
                ctx.spectral_radius_noalloc(d_res.value, d_out.value)
                ctx.synchronize()
                assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_out, 8, D2H) == 0
            finally:
                hip.hipFree(d_res)
                hip.hipFree(d_out)
        print(f"{cfg['q'].shape[:2]}: noalloc {out[0]:.17e} allocating {ref:.17e}")
        assert out[0] == ref
