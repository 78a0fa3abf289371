"""The NLTE ALI driver (dispatch_amd.ali) run on the host through the oracle's
restatements of every entry point (tests/ali_cases.OracleCtx): the orchestration
itself — windows, directions, the ALI step, Ng's acceleration — on a small box of
the workload generator."""
import numpy as np

from ali_cases import OracleCtx
from dispatch_amd.ali import AliConfig, NlteProblem


def small(**kw):
    base = dict(nx=6, ny=6, nz=48, nq=16, nbound=4, nline=6, ncont=4, lines_per_window=3, nphi=2, ncont_f=8)
    base.update(kw)
    return AliConfig(**base)


def test_converges_with_ng_and_departs_from_lte():
    p = NlteProblem(OracleCtx(), small(), device="cpu")
    n = p.solve()
    errs = [abs(e) for e, _ in p.history]
    assert errs[-1] < 1e-4 and n <= 40, errs
    assert np.abs(p.b.numpy() - 1).max() > 1e-2


def test_ng_shortens_the_iteration():
    q = NlteProblem(OracleCtx(), small(ng=False, max_iter=25), device="cpu")
    q.solve()
    p = NlteProblem(OracleCtx(), small(max_iter=25), device="cpu")
    p.solve()
    # after 25 iterations the accelerated run is an order of magnitude closer
    assert abs(p.history[-1][0]) < 0.1 * abs(q.history[-1][0]), (p.history[-1][0], q.history[-1][0])


def test_strong_collisions_give_lte():
    cfg = small(ne_scale=1e7)
    p = NlteProblem(OracleCtx(), cfg, device="cpu")
    p.solve()
    b = p.b.numpy().reshape(p.atom.nk, cfg.nz, -1)
    assert np.abs(b[:, :cfg.nz // 2] - 1).max() < 1e-3


def test_fixed_ng_schedule_follows_fortran_mod():
    """mod(iter - ng_start, ng_step) == 1 with Fortran's mod: with ng_step 4 and
    ng_start 8 the extrapolation fires at 9, 13, 17, ...; before ng_start the
    remainder is negative or zero and nothing fires (Python's % would fire at 1, 5)."""
    from dispatch_amd.ali import ng_fixed_fires
    assert [i for i in range(1, 30) if ng_fixed_fires(i, 4, 8)] == [9, 13, 17, 21, 25, 29]
    assert [i for i in range(1, 20) if ng_fixed_fires(i, 3, 6)] == [7, 10, 13, 16, 19]
    assert not any(ng_fixed_fires(i, 1, 8) for i in range(1, 40))   # mod(., 1) is never 1


def test_ng_step_below_three_is_raised_to_three():
    """as the reference's namelist check (SRC/m3d_patch_mod.f90:472-476): Ng needs 3
    iterates after an extrapolation, so ng_step 1 or 2 becomes 3 (firing at 9, 12, ...)"""
    import warnings
    from dispatch_amd.ali import AliConfig, ng_fixed_fires
    for step in (1, 2):
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            cfg = AliConfig(ng_step=step)
        assert cfg.ng_step == 3 and any("set to 3" in str(x.message) for x in w)
        assert [i for i in range(1, 20) if ng_fixed_fires(i, cfg.ng_step, cfg.ng_start)] == [9, 12, 15, 18]
    assert AliConfig(ng_step=0).ng_step == 0 and AliConfig(ng_step=4).ng_step == 4


def test_ng_gate_parameters_reach_the_gate():
    """The automatic gate's &nlte_params (SRC/m3d_patch_mod.f90:405-406) set through
    AliConfig.ng_gate: a contraction-rate floor no iteration reaches keeps Ng from
    firing at all, where the defaults fire; an unknown name is an error."""
    import pytest
    from dispatch_amd.ali import NgAcceleration
    p = NlteProblem(OracleCtx(), small(max_iter=15), device="cpu")
    p.solve()
    q = NlteProblem(OracleCtx(), small(max_iter=15, ng_gate={"rate_min": 0.9999}), device="cpu")
    q.solve()
    assert any(p.ng_history) and not any(q.ng_history)
    assert NgAcceleration(min_gap="3").min_gap == 3 and NgAcceleration.min_gap == 4
    with pytest.raises(ValueError, match="unknown Ng gate parameter"):
        NgAcceleration(gap=3)
