"""Writes eigen_3x3.npz: symmetric 3x3 matrices and their eigenvalues from
numpy.linalg.eigh (LAPACK syevd), the library routine the oracle's Jacobi
eigen-solve is checked against. Run from the repository root:

    python tests/golden/make_eigen_fixtures.py
"""
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "eigen_3x3.npz"


def rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.diag(r))


def main():
    rng = np.random.default_rng(20261203)
    mats = []
    # covariance-like: random orthonormal frame, eigenvalues over several decades
    for _ in range(300):
        lam = np.sort(10.0 ** rng.uniform(-6, 0, size=3))
        q = rotation(rng)
        mats.append(q @ np.diag(lam) @ q.T)
    # near-degenerate pairs (small eigen-gaps) and exact repeats
    for gap in (1e-3, 1e-5, 1e-7, 0.0):
        for _ in range(25):
            a = rng.uniform(0.1, 1.0)
            lam = np.array([rng.uniform(0.0, 0.05), a, a + gap * a])
            q = rotation(rng)
            mats.append(q @ np.diag(lam) @ q.T)
    # planar (one zero eigenvalue), linear (two zeros), isotropic
    for _ in range(25):
        q = rotation(rng)
        mats.append(q @ np.diag([0.0, rng.uniform(0.1, 1), rng.uniform(0.1, 1)]) @ q.T)
        mats.append(q @ np.diag([0.0, 0.0, rng.uniform(0.1, 1)]) @ q.T)
        mats.append(np.eye(3) * rng.uniform(0.1, 1))
    A = np.array(mats)
    A = 0.5 * (A + np.transpose(A, (0, 2, 1)))
    a6 = np.stack([A[:, 0, 0], A[:, 0, 1], A[:, 0, 2], A[:, 1, 1], A[:, 1, 2], A[:, 2, 2]], axis=1)
    ev = np.linalg.eigh(A)[0]
    np.savez(OUT, a6=a6, ev=ev)
    print(f"wrote {OUT} ({len(a6)} matrices)")


if __name__ == "__main__":
    main()
