"""Pins the oracle's direction post-processing (direction/src/pca.cu:67-106)
by properties the formulas fix, not by restating them:

* calcAngleFromMD: the angles rebuild the direction.  (sin t cos p, sin t sin p,
  cos t) equals md / |md| when md_y >= 0 and -md / |md| when md_y < 0 (the
  reference flips the direction into the y >= 0 half space), so phi lies in
  [0, pi] and theta in [0, pi]; axis cases give their textbook angles.
* directionNormalization: |md| > 0.65 gives a unit vector along md, anything
  else (the threshold itself included) gives 0.
"""
import numpy as np


def test_angles_rebuild_the_direction(orc):
    rng = np.random.default_rng(70)
    md = rng.standard_normal((3, 5000)).astype(np.float32) * rng.random(5000).astype(np.float32)
    phi, theta = orc.md_angles(md)
    assert (phi >= 0).all() and (phi <= np.pi + 1e-6).all() and (theta >= 0).all() and (theta <= np.pi + 1e-6).all()
    u = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    flip = np.where(md[1] < 0, -1.0, 1.0)
    want = flip * md / np.linalg.norm(md.astype(np.float64), axis=0)
    assert np.abs(u - want).max() < 1e-6


def test_angles_axis_cases(orc):
    md = np.array([[1, 0, 0, 0, -1, 0], [0, 1, 0, -1, 0, 0], [0, 0, 1, 0, 0, -1]], np.float32)
    phi, theta = orc.md_angles(md)
    want_phi = [0, np.pi / 2, 0, np.pi / 2, np.pi, 0]
    want_theta = [np.pi / 2, np.pi / 2, 0, np.pi / 2, np.pi / 2, np.pi]
    assert np.allclose(phi, want_phi, atol=1e-7) and np.allclose(theta, want_theta, atol=1e-7)


def test_normalisation_threshold(orc):
    rng = np.random.default_rng(71)
    md = (rng.standard_normal((3, 4000)) * 0.5).astype(np.float32)
    md[:, 0] = [0.65, 0, 0]  # exactly the threshold: not above it
    md[:, 1] = [0, 0.6500001, 0]
    out = orc.md_normalise(md)
    n = np.sqrt((md.astype(np.float64) ** 2).sum(axis=0))
    big = n > 0.65
    assert np.allclose(np.linalg.norm(out[:, big], axis=0), 1.0, atol=1e-6)
    assert np.allclose(out[:, big], md[:, big] / n[big], atol=1e-6)
    assert (out[:, ~big] == 0).all()
    assert (out[:, 0] == 0).all() and out[1, 1] == 1.0
