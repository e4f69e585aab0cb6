"""HIP-graph capture of device-pointer applies (tcx_graph_*).

Every replay is compared with the eager result of the same calls, with eager
work and synchronisations between the replays; the outputs are cleared inside
the graph by the library's fill kernel (no hipMemsetAsync node)."""
import numpy as np
import pytest

from tc_amd import system as S
from tc_amd import tcx

pytestmark = pytest.mark.gpu


def _setup(engine):
    sys_ = S.tiny_system(method="TC", n=(12, 12, 12), kmesh=(2, 2, 1), nbands=4, nocc=3, seed_stream=71)
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[4, 2, 3, 1], seed=72)
    return sys_, nb, phi


def test_graph_replays_match_eager(engine):
    sys_, nb, phi = _setup(engine)
    d_phi = engine.to_device(phi)
    d_out = engine.device_buffer(phi.nbytes)
    d_side = engine.device_buffer(phi.nbytes)

    def calls():
        d_out.zero()
        engine._term_device("hf2x", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP32)
        engine._term_device("tc2x", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP32)

    calls()
    engine.synchronize()
    ref = d_out.download(np.complex128)
    assert np.abs(ref).max() > 0
    engine.graph_begin()
    calls()
    gid = engine.graph_end()
    for rep in range(4):
        d_out.upload(np.full(phi.size, 7.0 + 1j))      # garbage the graph must clear
        engine.graph_launch(gid)
        engine.synchronize()
        got = d_out.download(np.complex128)
        print("replay", rep, np.abs(got - ref).max(), flush=True)
        assert np.array_equal(got, ref)
        d_side.zero()                                    # eager work between replays
        engine._term_device("hf2x", nb, d_phi.ptr, d_side.ptr, tcx.PREC_FP64)
        engine.synchronize()
    engine.graph_destroy(gid)
    for b in (d_phi, d_out, d_side):
        b.free()


def test_graph_capture_errors_leave_the_context_usable(engine):
    sys_, nb, phi = _setup(engine)
    d_phi = engine.to_device(phi)
    d_out = engine.device_buffer(phi.nbytes)
    engine.graph_begin()
    with pytest.raises(tcx.TcxError):
        engine.hf2x(nb, phi)                             # host pointers cannot be captured
    engine.graph_begin()
    with pytest.raises(tcx.TcxError):                    # a band layout never run before
        engine._term_device("hf2x", [1, 1, 1, 1], d_phi.ptr, d_out.ptr, tcx.PREC_FP32)
    with pytest.raises(tcx.TcxError):
        engine.graph_end()                               # the failed call dropped the capture
    d_out.zero()
    engine._term_device("hf2x", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP64)
    got = d_out.download(np.complex128)
    assert np.allclose(got, engine.hf2x(nb, phi), rtol=0, atol=1e-12 * np.abs(got).max())
    d_phi.free()
    d_out.free()
