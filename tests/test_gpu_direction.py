"""Direction post-processing on the GPU (xr_md_angles / xr_md_normalise and
their host-buffer variants) against the oracle's restatement of
direction/src/pca.cu:67-106, the products and sums rounded as the
reference's host code rounds them (no FMA contraction): normalisation is
exact (correctly rounded sqrt and division), the angles lie within 4 ulp
(the device atan2f)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def sample_md(n, seed):
    rng = np.random.default_rng(seed)
    md = (rng.standard_normal((3, n)) * rng.random(n)).astype(np.float32)
    md[:, :8] = np.array([[0.65, 0, 0], [0, 0.6500001, 0], [0, 0, 0], [1, 0, 0], [0, -1, 0], [0, 0, -1],
                          [-1e-20, 0, 0], [0, 0, 1]], np.float32).T
    return md


def test_md_angles_match_oracle(orc, xr, ctx, torch_dev):
    import torch

    n = 300001
    md = sample_md(n, 80)
    want = orc.md_angles(md)
    got_host = ctx.md_angles_host(md)
    dmd = torch.from_numpy(md.reshape(-1)).to(torch_dev)
    dang = torch.empty(2 * n, device=torch_dev)
    ctx.md_angles(dmd, dang, n)
    got_dev = dang.cpu().numpy().reshape(2, n)
    assert np.array_equal(got_host, got_dev)
    # atan2f: the device library and the host's differ by a few ulp (HIP
    # documents 2 ulp for atan2f; glibc's is within 1)
    ulp = np.spacing(np.maximum(np.abs(want), np.float32(1.0)))
    assert (np.abs(got_dev - want) <= 4 * ulp).all()


def test_md_normalise_matches_oracle(orc, xr, ctx, torch_dev):
    import torch

    n = 300001
    md = sample_md(n, 81)
    want = orc.md_normalise(md)
    got_host = ctx.md_normalise_host(md)
    dmd = torch.from_numpy(md.reshape(-1)).to(torch_dev)
    ctx.md_normalise(dmd, n)
    got_dev = dmd.cpu().numpy().reshape(3, n)
    assert np.array_equal(got_host, got_dev)
    assert np.array_equal(got_dev, want)
