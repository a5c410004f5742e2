"""Writes svd_vectors.npz: golden 3x3 SVD / polar / clamp vectors computed
with numpy and scipy (independent library routines), used to pin the
oracle's Jacobi SVD for the non-diagonal and inverted inputs that the
reference's own tests (all diagonal) leave open.

Rows: F (random, near-identity, rotated, inverted det<0, repeated singular
values); S = numpy.linalg.svd singular values; R = scipy.linalg.polar
orthogonal factor; C = U clamp(S, 1-0.025, 1+0.0075) V^T built from numpy's
U, V (a matrix function of F, unique for distinct or repeated S).
Run: python tests/golden/make_svd_vectors.py
"""
import os

import numpy as np
import scipy.linalg

rng = np.random.default_rng(20261203)
Fs = []
for _ in range(40):                      # general random
    Fs.append(rng.normal(size=(3, 3)))
for _ in range(40):                      # near identity (the normal MPM case)
    Fs.append(np.eye(3) + 0.03 * rng.normal(size=(3, 3)))
for _ in range(20):                      # rotation times stretch
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    Fs.append(Q @ np.diag(1 + 0.05 * rng.normal(size=3)))
for _ in range(20):                      # inverted (det < 0)
    F = np.eye(3) + 0.1 * rng.normal(size=(3, 3))
    F[:, 1] *= -1
    Fs.append(F)
for _ in range(10):                      # repeated singular values
    Q1, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Q2, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    s = 0.9 + 0.2 * rng.random()
    Fs.append(Q1 @ np.diag([s, s, 1.1 * s]) @ Q2.T)
F = np.array(Fs)
S = np.zeros((len(F), 3))
R = np.zeros_like(F)
Cl = np.zeros_like(F)
for k, A in enumerate(F):
    U, s, Vt = np.linalg.svd(A)
    S[k] = s
    R[k] = scipy.linalg.polar(A)[0]
    Cl[k] = U @ np.diag(np.clip(s, 1 - 0.025, 1 + 0.0075)) @ Vt
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "svd_vectors.npz"),
                    F=F, S=S, R=R, C=Cl)
print("wrote", len(F), "vectors")
