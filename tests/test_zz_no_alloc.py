"""tcx_hf2x_noalloc, the exchange apply that allocates no device memory, against tcx_hf2x.

The entry point takes its working memory from a fixed device array (256 MiB per GPU)
and walks the applied bands in chunks that fit it, with the same kernels as tcx_hf2x.
Every pair record is computed identically; only the order of the FP64 band-pair sums
differs (other source blocks per launch), so both precisions agree with tcx_hf2x to
1e-12 relative (max |delta| / max |reference|).  Checked on two calls in a row, the
second with fewer bands, on a cell whose 64 applied bands need several chunks; plus the
entry point's errors: occupied grids not yet built for the precision, host pointers.
"""
import ctypes as C

import numpy as np
import pytest

from tc_amd import system as S
from tc_amd import tcx

pytestmark = pytest.mark.gpu


def _noalloc(eng, nb, phi_ptr, out_ptr, prec):
    nb = np.ascontiguousarray(nb, dtype=np.int32)
    return eng.L.tcx_hf2x_noalloc(eng.h, nb.ctypes.data_as(C.POINTER(C.c_int)), C.c_void_p(phi_ptr),
                                  C.c_void_p(out_ptr), int(prec))


@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32], ids=["fp64", "fp32"])
def test_hf2x_noalloc_matches_hf2x(prec):
    sys_ = S.tiny_system(method="HF", n=(24, 24, 24), kmesh=(2, 2, 2), nbands=8, nocc=7, seed_stream=81)
    eng = tcx.Engine(0)
    bufs = []
    try:
        eng.load_system(sys_)
        nb0 = np.full(sys_.nk, 8, np.int32)
        _, phi0 = S.applied_bands(sys_, nb=nb0, seed=82)
        d_phi = eng.to_device(phi0)
        d_out = eng.device_buffer(phi0.nbytes)
        bufs += [d_phi, d_out]
        # before any apply in this precision the occupied grids are not built: refused
        st = _noalloc(eng, nb0, d_phi.ptr, d_out.ptr, prec)
        print(f"noalloc before the occupied grids: status {st} ({eng.L.tcx_last_error(eng.h)})", flush=True)
        assert st == tcx.TCX_ERR_STATE
        # host pointers are refused (the entry point takes device memory only)
        host_out = np.zeros_like(phi0)
        eng._term_device("hf2x", nb0, d_phi.ptr, d_out.ptr, prec)  # builds the occupied grids
        st = _noalloc(eng, nb0, phi0.ctypes.data, host_out.ctypes.data, prec)
        assert st == tcx.TCX_ERR_ARG
        for call, nb in enumerate((nb0, np.array([3, 0, 2, 1, 0, 4, 1, 2], np.int32))):
            _, phi = S.applied_bands(sys_, nb=nb, seed=83 + call)
            d_in = eng.to_device(phi)
            d_ref = eng.device_buffer(phi.nbytes)
            d_got = eng.device_buffer(phi.nbytes)
            bufs += [d_in, d_ref, d_got]
            d_ref.zero()
            eng._term_device("hf2x", nb, d_in.ptr, d_ref.ptr, prec)
            d_got.upload(np.full(phi.size, 0.25 - 0.5j))  # the entry point accumulates into its output
            print(f"noalloc call {call}: {int(nb.sum())} bands, precision {prec}", flush=True)
            st = _noalloc(eng, nb, d_in.ptr, d_got.ptr, prec)
            assert st == tcx.TCX_OK, eng.L.tcx_last_error(eng.h)
            ref = d_ref.download(np.complex128)
            got = d_got.download(np.complex128) - (0.25 - 0.5j)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"noalloc call {call}: max rel {err:.3e}", flush=True)
            assert err <= 1e-12
    finally:
        for b in bufs:
            b.free()
        eng.close()
