"""GPU: a batch advances by exactly the steps asked for.  shll_batch_advance(n) with n not
a multiple of the launch's steps ends in a shorter launch; graph chunks that follow start
at any step.  After every call each instance equals its own oracle run at exactly
min(steps taken, its step count) (bitwise; F64F within round-off, as in test_gpu_parity);
the F64F and C32K batches are also checked at the launch
shapes the benchmark times (8 cells per lane, 8 or 16 steps per launch, one wavefront per block,
128-step graphs) with multi-segment tubes and full graph chunks."""
import pytest

import shll_amd as S
from test_gpu_parity import assert_oracle_match

pytestmark = pytest.mark.gpu

MODES = {S.F64: 0, S.REF32: 1, S.C32: 2, S.C32K: 3, S.F64F: 4}


def banner(*a):
    print("[gpu-case]", *a, flush=True)


def opb(orc, pb):
    return orc.problem(pb.dim, pb.order, pb.nx, pb.ny, bc=pb.bc, lx=pb.lx, ly=pb.ly, limiter=pb.limiter,
                       theta=pb.theta)


def check_all(orc, b, pbs, prims, steps, done, what):
    for k in range(len(pbs)):
        n = min(done, steps[k])
        _, ref = orc.run(opb(orc, pbs[k]), MODES[pbs[k].precision], prims[k], n, nthreads=4)
        assert_oracle_match(pbs[k].precision, b.conserved(k), ref, "%s: instance %d after %d steps" % (what, k, n))


@pytest.mark.parametrize("prec", [S.F64, S.C32K])
@pytest.mark.parametrize("dim", [1, 2])
def test_advance_takes_exactly_the_steps_asked_for(orc, prec, dim):
    if dim == 1:
        pbs, prims, steps = S.tube_batch(2468, 7, nmin=100, nmax=2000, precision=prec)
        steps = [max(1, s // 10) for s in steps]
        launch = (2, 8, 4, 32)   # 4 launches of 8 steps per graph chunk
    else:
        pbs, prims = S.domain_batch(135, 3, smin=30, smax=80, precision=prec)
        steps = [70, 23, 51]
        launch = (1, 2 if prec == S.F64 else 4, 4, 16)
    b = S.Batch(pbs, prims, steps)
    b.set_launch(*launch)
    k = b.get_launch()[1]
    total = max(steps)
    done = 0
    # pieces that leave every remainder modulo k, with graph chunks anchored at odd steps
    for n in (5, 45, 3, 1, 38, 2 * k + 1):
        banner("exact advance", prec, dim, launch, n)
        rem = b.advance(n)
        done = min(total, done + n)
        assert rem == total - done
        assert b.last_timing()[1] >= 0
        check_all(orc, b, pbs, prims, steps, done, "dim %d piece %d" % (dim, n))
    b.run()
    check_all(orc, b, pbs, prims, steps, total, "dim %d run" % dim)
    b.close()


@pytest.mark.parametrize("prec,launch", [(S.F64F, (8, 8, 1, 128)), (S.C32K, (8, 16, 1, 128))])
def test_headline_launch_shape_batch(orc, prec, launch):
    """Tubes of 4096-9000 cells (several 512-cell segments each) at the benchmark's launch
    shapes (F64F, the timed form: 8 steps per launch, 16-cell halos; C32K: 16 steps, 32-cell
    halos): a 20-step advance (a shorter last launch), then 128-step graph chunks from step 20
    on, instances finishing inside chunks."""
    pbs, prims, steps = S.tube_batch(97531, 6, nmin=4096, nmax=9000, precision=prec)
    steps = [310, 20, 251, 137, 300, 129]
    b = S.Batch(pbs, prims, steps)
    b.set_launch(*launch)
    banner("headline shape batch", prec, launch, [p.nx for p in pbs])
    assert b.advance(20) == 290
    check_all(orc, b, pbs, prims, steps, 20, "after 20 steps")
    b.run()
    check_all(orc, b, pbs, prims, steps, 310, "whole run")
    # and once more from the start in one call (chunks anchored at step 0)
    b.reset()
    b.run()
    check_all(orc, b, pbs, prims, steps, 310, "whole run after reset")
    b.close()
