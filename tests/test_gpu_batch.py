"""Batched uneven slices (128^2, 256^2, 384^2; 8 coils): one persistent F^H d
launch, per-class Q, autotuned schedules, graph-captured PCG per slice."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_batch_slices_converge_and_match_single_slice_path(ctx):
    from impatient import batch, ops, IMP_CG_BEST_ITERATE, IMP_CG_PRECOND, IMP_CG_TRUE_RESIDUAL
    from oracle import oracle as orc
    sl = batch.make_slices(6, seed=77, ctx=ctx)
    sizes = sorted({s.prob.Nx for s in sl})
    res = batch.reconstruct(ctx, sl, tol=1e-5)
    recs = res["slices"]
    print({k: v for k, v in res.items() if k != "slices"})
    for r in recs:
        print(f"N={r['N']}: {r['iterations']} it, true {r['true_rel_residual']:.2e}, "
              f"{r['slice_s'] * 1e3:.1f} ms, cpg {r['cpg']}")
    assert all(r["converged"] == 1 for r in recs)
    assert max(r["true_rel_residual"] for r in recs) < 1e-5
    assert len(sizes) >= 2
    # the first slice of each size through the single-image operators
    for N in sizes:
        i = next(k for k, s in enumerate(sl) if s.prob.Nx == N)
        p = sl[i].prob
        b = ops.nudft_adjoint(ctx, p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, sl[i].d, p.sens)
        q = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, N, N, 1, p.ntime_segs)
        tau, t0 = orc.segment_timing(p.t, p.ntime_segs)
        op = ops.Toeplitz(ctx, p.shape, q, tau, t0, p.fm, p.sens, p.lam)
        x, st = op.cg(b, 500, 1e-5, IMP_CG_BEST_ITERATE | IMP_CG_TRUE_RESIDUAL | IMP_CG_PRECOND)
        print(f"N={N}: single path {st['iterations']} it vs batch {recs[i]['iterations']} it")
        assert abs(st["iterations"] - recs[i]["iterations"]) <= 2
        op.close()
