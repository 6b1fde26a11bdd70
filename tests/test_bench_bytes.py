"""The benchmark's algorithmic-bytes formula B_mv against the element counts
worked out independently for the configurations (sphere 40 and 128 across,
the 128-RHS ellipsoid block) and against the geometry the benchmark builds."""
import bench
from dda_amd import problem as P


def test_b_mv_element_counts():
    # sphere 40 across: nv 33,552, box 40^3, grid 80^3
    assert bench.b_mv(33552, (40, 40, 40), (80, 80, 80), 1) == 7_981_968
    # sphere 128 across: nv 1,099,136, box 128^3, grid 256^3 (FP64 4.185 GB, FP32 2.092 GB)
    assert bench.b_mv(1099136, (128, 128, 128), (256, 256, 256), 1) == 261_550_464
    assert bench.b_mv(1099136, (128, 128, 128), (256, 256, 256), c=16) == 4_184_807_424
    assert bench.b_mv(1099136, (128, 128, 128), (256, 256, 256), c=8) == 2_092_403_712
    # ellipsoid 32x24x16 semi-axes: nv 51,488, box 64x48x32, grid 128x96x64; one RHS and a
    # 128-RHS block sharing the tensor
    assert bench.b_mv(51488, (64, 48, 32), (128, 96, 64), 1) == 12_259_872
    assert bench.b_mv(51488, (64, 48, 32), (128, 96, 64), 1, nrhs=128) == 970_002_432


def test_headline_geometry_matches_the_counts():
    g = P.sphere(64.0, 1.5 + 0.1j, 1e-7)
    params = P.DDAParams(g, 2 * 3.141592653589793 * 1e-7 * abs(1.5 + 0.1j) / 0.5)
    assert params.nv == 1099136 and tuple(g.box) == (128, 128, 128) and tuple(params.grid) == (256, 256, 256)
    assert params.vec_stride == 1099136  # already a multiple of 16
