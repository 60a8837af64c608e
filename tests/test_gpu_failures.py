"""Error paths forced from the host: device allocations made to fail by the
library's test hook (xr_test_fail_allocations) at every point of a solve.
Each entry point must return XR_ERR_HIP with the allocation named in
xr_last_error, release what it had allocated, and leave the library usable:
the same solve run afterwards gives the result of an undisturbed run."""
import numpy as np
import pytest

from conftest import ellipsoid_phantom
from test_gpu_parity import Setup

pytestmark = pytest.mark.gpu


@pytest.fixture
def hook(xr):
    yield xr.test_fail_allocations
    xr.test_fail_allocations(-1)


def test_context_creation_failure_is_clean(xr, hook):
    hook(0)
    with pytest.raises(xr.XRError):
        xr.Context(0)
    hook(-1)
    c = xr.Context(0)
    c.close()


def test_solver_allocation_failures(orc, xr, ctx, hook):
    st = Setup(orc, xr, N=24, U=40, P=12, vs=0.6)
    truth, _ = ellipsoid_phantom(orc, xr, st.N, st.vs, seed=70)
    sino = np.stack([st.case.forward_view(truth, n)[0] for n in range(st.P)]).astype(np.float32).reshape(-1)
    vol0 = np.full(st.N ** 3, 0.01, np.float32)

    def solve():
        v = [vol0.copy()]
        losses = ctx.ir_reconstruct(st.geom, st.scan, [sino.copy()], v, 2, 3)
        return v[0], losses

    want_v, want_l = solve()
    failures = 0
    for after in range(0, 40):
        hook(after)
        try:
            solve()
            hook(-1)
            break  # the solve needs fewer allocations: every failure point was covered
        except xr.XRError as e:
            failures += 1
            # names the failed allocation (a hipMalloc call or a device buffer's alloc)
            assert "out of memory" in str(e) and "alloc" in str(e), str(e)
        hook(-1)
        got_v, got_l = solve()  # the library recovers
        # the volume bit for bit; the FP64 loss sums to rounding (atomic order)
        assert np.array_equal(got_v, want_v) and np.allclose(got_l, want_l, rtol=1e-12, atol=0)
    assert failures >= 3


def test_plan_and_batch_allocation_failures(orc, xr, ctx, hook):
    st = Setup(orc, xr, N=20, U=32, P=8, vs=0.6)
    hook(0)
    with pytest.raises(xr.XRError):
        ctx.plan(st.geom, st.scan)
    with pytest.raises(xr.XRError):
        ctx.batch([st.geom, st.geom])
    hook(-1)
    plan = ctx.plan(st.geom, st.scan)
    plan.close()
    b = ctx.batch([st.geom, st.geom])
    b.close()


def test_batch_run_allocation_failures(orc, xr, ctx, hook):
    """A batch whose run fails while growing its subset lists or building its
    tables runs normally afterwards, with the result of an undisturbed run."""
    import torch

    st = Setup(orc, xr, N=20, U=32, P=8, vs=0.6)
    truth, _ = ellipsoid_phantom(orc, xr, st.N, st.vs, seed=71)
    sino = torch.from_numpy(np.stack([st.case.forward_view(truth, n)[0] for n in range(st.P)])
                            .astype(np.float32).reshape(-1)).cuda()

    def run(b, epochs):
        for i in range(2):
            va, sa = b.buffers(i)
            ctx.copy(sa, sino)
            ctx.copy(va, torch.full((st.N ** 3,), 0.01, device="cuda"))
        b.restart()
        return b.run(epochs)

    ref = ctx.batch([st.geom, st.geom], flags=0)
    want = run(ref, 3)
    ref.close()
    failures = 0
    for after in range(0, 30):
        b = ctx.batch([st.geom, st.geom], flags=0)
        hook(after)
        try:
            run(b, 3)
            hook(-1)
            b.close()
            break
        except xr.XRError:
            failures += 1
        hook(-1)
        got = run(b, 3)
        assert np.allclose(got, want, rtol=1e-12, atol=0)  # FP64 loss sums, atomic order
        b.close()
    assert failures >= 2
