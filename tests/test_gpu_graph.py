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


def test_mixed_grid_graphs_vs_oracle(engine):
    """Three contexts at different FFT grids (8^3 HF, 12^3 TC, 16^3 BITC), each with a graph
    of FP32-split device applies, replayed in turn with eager work between the rounds. Every
    replay is checked against the FP64 oracle (per band ||delta|| / ||ref|| <= 3e-5, the
    FP32-split bound), not only against the eager result."""
    from oracle import oracle as O
    from test_gpu_hf2x import per_band_rel

    cells = [("HF", (8, 8, 8), ("hf2x",)), ("TC", (12, 12, 12), ("hf2x", "tc2x")),
             ("BITC", (16, 16, 16), ("tc2x",))]
    engines = [engine] + [tcx.Engine(0) for _ in cells[1:]]
    runs = []
    try:
        for i, ((method, n, terms), eng) in enumerate(zip(cells, engines)):
            sys_ = S.tiny_system(method=method, n=n, kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=81 + i)
            eng.load_system(sys_)
            nb, phi = S.applied_bands(sys_, nb=[3, 1], seed=91 + i)
            o = O.Oracle(sys_)
            ref = sum(getattr(o, t)(nb, phi) for t in terms)
            d_phi = eng.to_device(phi)
            d_out = eng.device_buffer(phi.nbytes)

            def calls(eng=eng, nb=nb, d_phi=d_phi, d_out=d_out, terms=terms):
                d_out.zero()
                for t in terms:
                    eng._term_device(t, nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP32)

            calls()                                      # first run builds the plans
            eng.synchronize()
            eng.graph_begin()
            calls()
            gid = eng.graph_end()
            runs.append((sys_, nb, ref, eng, gid, d_phi, d_out))
        for rep in range(3):
            for sys_, nb, ref, eng, gid, d_phi, d_out in runs:
                d_out.upload(np.full(ref.size, 3.0 - 2j))   # garbage the graph must clear
                eng.graph_launch(gid)
            for sys_, nb, ref, eng, gid, d_phi, d_out in runs:
                eng.synchronize()
                got = d_out.download(np.complex128)
                rel = per_band_rel(sys_, nb, got, ref)
                print(f"replay {rep} {sys_.n} per-band rel max {rel.max():.3e}", flush=True)
                assert rel.max() <= 3e-5, (rep, sys_.n)
                eng._term_device("hf2x", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP64)  # eager work between rounds
                eng.synchronize()
    finally:
        for sys_, nb, ref, eng, gid, d_phi, d_out in runs:
            eng.graph_destroy(gid)
            d_phi.free()
            d_out.free()
        for eng in engines[1:]:
            eng.close()
