"""The two pass orders of the FP32-split exchange pipeline against the oracle and each other.

Plane-first (xy planes, z pencils, xy planes; selected with tcx_set_pipeline_order) keeps
only the (kx, ky) columns of the pair records inside the pair density's G-space support; z-first keeps them
all.  Both are checked at every compiled grid size against the FP64 oracle (per band
||delta|| / ||ref|| <= 3e-5, the FP32-split bound) and against each other (per band
<= 1e-6: the same FP32 transforms in another order, FP64 sums), with ragged band counts and
several launch shapes, and the plane-first records must hold fewer columns than the grid.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tc_amd import system as S
from tc_amd import tcx
from test_gpu_hf2x import per_band_rel

pytestmark = pytest.mark.gpu


def cell(grid, seed):
    if grid >= 32:  # the benchmark's fcc cell, reduced k-mesh and bands
        return S.build_system(S.fcc_lattice(10.26), (grid,) * 3, (2, 1, 1), nbands=3, nocc=2, seed_stream=seed,
                              name=f"HF fcc a10.26 2x1x1k {grid}^3")
    return S.tiny_system(method="HF", n=(grid,) * 3, kmesh=(2, 2, 1), nbands=3, nocc=2, seed_stream=seed)


@pytest.mark.parametrize("grid", [6, 8, 12, 16, 20, 24, 32, 36, 40, 48])
def test_plane_first_and_z_first_vs_oracle(engine, grid):
    sys_ = cell(grid, 500 + grid)
    engine.load_system(sys_)
    nb = np.array([(i % 3) for i in range(sys_.nk)], np.int32)  # ragged, with an empty k
    _, phi = S.applied_bands(sys_, nb=nb, seed=501)
    ref = O.Oracle(sys_).hf2x(nb, phi)
    out = {}
    for order in (1, 0):
        engine.set_pipeline_order("hf2x", order)
        out[order] = engine.hf2x(nb, phi, precision=tcx.PREC_FP32)
        cols = engine.record_columns("hf2x")
        rel = per_band_rel(sys_, nb, out[order], ref)
        print(f"{grid}^3 order {order}: per-band rel max {rel.max():.3e}, record columns {cols:.1f} of {grid * grid}",
              flush=True)
        assert rel.max() <= 3e-5
        if order == 1:
            assert cols < grid * grid
        else:
            assert cols == grid * grid
    engine.set_pipeline_order("hf2x", -1)
    rel = per_band_rel(sys_, nb, out[1], out[0])
    print(f"{grid}^3 plane-first vs z-first per-band rel max {rel.max():.3e}", flush=True)
    assert rel.max() <= 1e-6


@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (7, 5), (0, 0)])
def test_plane_first_launch_shapes(engine, shape):
    """Results do not depend on how pairs are grouped into launches (FP64 sums in another
    order: 1e-12)."""
    sys_ = cell(16, 520)
    engine.load_system(sys_)
    nb = np.array([2, 0, 1, 2], np.int32)
    _, phi = S.applied_bands(sys_, nb=nb, seed=521)
    engine.set_pipeline_order("hf2x", 1)
    engine.set_launch_shape("hf2x", 0, 0)
    base = engine.hf2x(nb, phi, precision=tcx.PREC_FP32)
    engine.set_launch_shape("hf2x", *shape)
    got = engine.hf2x(nb, phi, precision=tcx.PREC_FP32)
    engine.set_launch_shape("hf2x", 0, 0)
    engine.set_pipeline_order("hf2x", -1)
    err = np.abs(got - base).max() / np.abs(base).max()
    print("shape", shape, err, flush=True)
    assert err <= 1e-12


def test_pipeline_order_errors(engine):
    for term, order in ((1, 0), (0, 2), (0, -2)):
        with pytest.raises(tcx.TcxError) as e:
            engine.set_pipeline_order(term, order)
        assert e.value.code == tcx.TCX_ERR_ARG
