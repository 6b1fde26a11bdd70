"""Per-pass event timing (dda_op_profile / dda_op_profile_read).

Every launch of a plain matvec is a full launch; in a device-side solve the
launches queued past a right-hand side's convergence return at entry and are
counted as skipped, so the per-pass averages are those of full launches."""
import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu


def test_profile_counts_full_matvecs(ctx):
    g, params, cc, cs, prop, pol = cases.build("sphere_r7", grid=(32, 32, 32))
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=32, grid=params.grid)
    op.set_couple_constants(cs)
    x, y = D.DeviceVector(ctx, params.vec_size, 32), D.DeviceVector(ctx, params.vec_size, 32)
    x.upload(P.random_dipole_vector(params))
    op.apply(x, y)
    ctx.synchronize()
    op.profile(True)
    for _ in range(10):
        op.apply(x, y)
    ms, counts, skipped = op.profile_read()
    op.profile(False)
    assert skipped == 0 and list(counts) == [10] * 5
    assert np.all(ms > 0)
    op.close()


def test_profile_skips_launches_after_convergence(ctx):
    """A batched solve whose right-hand sides converge at different
    iterations: the passes keep (nearly) equal full-launch counts and the no-op
    launches queued past convergence are reported as skipped (a grid whose
    launches are long against a no-op one: 96^3)."""
    g, params, cc, cs, prop, pol = cases.build("sphere_c1", grid=(96, 96, 96))
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=64, grid=params.grid)
    _, pol1, pol2 = P.beam_pols(prop)
    fields = [P.einc_plane(params, prop, p) for p in (pol1, pol2)]
    op.set_couple_constants_batch(np.array([P.couple_constants(params, prop, p, "ldr")[1] for p in (pol1, pol2)]))
    s = D.GpuBicgStab(op)
    s.get_pol_vecs(np.array(fields), 1e-10)
    op.profile(True)
    _, infos = s.get_pol_vecs(np.array(fields), 1e-10)
    ms, counts, skipped = s.profile_read()
    op.profile(False)
    assert all(i["status"] == D.OK for i in infos)
    # a launch for part of the right-hand sides may land on either side of the
    # full-launch threshold in different passes
    assert counts.max() - counts.min() <= 2 and counts.min() > 0
    # two matvecs per iteration plus the start and the final true-residual
    # matvec: full launches while both right-hand sides run, none counted for
    # the no-op ones queued past convergence. A pass's events also enclose any
    # wait of the GPU for the host's next launch (at a chunk boundary), which can
    # lift a rare no-op launch over the threshold: two launches of slack.
    its = [i["iterations"] for i in infos]
    assert 2 * min(its) <= counts[0] <= 2 * max(its) + 4
    s.close()
    op.close()
