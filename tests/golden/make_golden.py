"""Writes the committed golden fixtures of tests/golden/.

eig3_numpy.npz: symmetric 3x3 matrices (random SPD, rank-deficient, repeated
and diagonal cases) with numpy.linalg.eigh eigenvalues / eigenvectors — the
textbook routine that pins the 3x3 eigen solver of the PCA step
(calcEigenVector uses Eigen's SelfAdjointEigenSolver, direction/src/pca.cu:33).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def eig_cases():
    rng = np.random.default_rng(20240601)
    mats, degen = [], []
    for _ in range(24):
        M = rng.normal(size=(3, 3))
        mats.append(M @ M.T)
        degen.append(False)
    for _ in range(6):  # rank 1 and rank 2 (PCA covariance of a few directions)
        v = rng.normal(size=(3, 2))
        mats.append(v @ v.T)
        degen.append(False)
        u = rng.normal(size=3)
        mats.append(np.outer(u, u))
        degen.append(True)  # two zero eigenvalues
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    mats.append(Q @ np.diag([1.0, 1.0, 3.0]) @ Q.T)
    degen.append(True)
    mats.append(np.diag([0.5, 0.2, 0.9]))
    degen.append(False)
    mats.append(np.zeros((3, 3)))
    degen.append(True)
    A = np.array(mats)
    w = np.zeros((len(A), 3))
    V = np.zeros((len(A), 3, 3))
    for i, M in enumerate(A):
        w[i], V[i] = np.linalg.eigh(M)
    return A, w, V, np.array(degen)


def main():
    A, w, V, degen = eig_cases()
    np.savez(os.path.join(HERE, "eig3_numpy.npz"), A=A, w=w, V=V, degenerate=degen)
    print("wrote", os.path.join(HERE, "eig3_numpy.npz"))


if __name__ == "__main__":
    main()
