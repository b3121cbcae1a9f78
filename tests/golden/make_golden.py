"""Writes tests/golden/closed_forms.npz: small input/output vectors computed with
numpy from the closed forms of the reference kernels, independent of the oracle.

  bh3  phi(r) = -slope * sqrt(r^2 + c^2)    (rbf/polyharmonic_odd.hpp:32-43, kSign=-1)
  th3  phi(r) = +slope * (r^2 + c^2)^(3/2)  (rbf/polyharmonic_odd.hpp:120-136)
  exp  phi(r) = psill * exp(-3 r / range)   (rbf/cov_exponential.hpp:33-40)
  anisotropy: d -> A d                        (rbf/rbf_base.hpp:40-43)
  monomials [1, x, y, z, x^2, xy, xz, y^2, yz, z^2] (polynomial/monomial_basis.hpp:198-247)

Run: python tests/golden/make_golden.py   (deterministic; seed 20240601)
"""
import os

import numpy as np

rng = np.random.default_rng(20240601)
out = {}

# kernel values on a few difference vectors, with and without anisotropy
d = rng.uniform(-2, 2, size=(16, 3))
d[0] = 0.0
aniso = np.array([[1.2, 0.1, 0.0], [0.0, 0.9, -0.2], [0.05, 0.0, 0.7]])
out["d"] = d
out["aniso"] = aniso
for name, a in (("iso", np.eye(3)), ("aniso", aniso)):
    ad = d @ a.T
    r = np.linalg.norm(ad, axis=1)
    out[f"bh3_{name}"] = -1.0 * np.sqrt(r**2 + 0.0)
    out[f"bh3c_{name}"] = -2.5 * np.sqrt(r**2 + 0.3**2)
    out[f"th3_{name}"] = 1.0 * r**3
    out[f"th3c_{name}"] = 0.7 * (r**2 + 0.2**2) ** 1.5
    out[f"exp_{name}"] = 1.0 * np.exp(-3.0 * r / 2.0)
    out[f"exp2_{name}"] = 0.4 * np.exp(-3.0 * r / 0.6)

# monomials
pts = rng.uniform(-1, 1, size=(7, 3))
out["mono_pts"] = pts
x, y, z = pts.T
full = np.stack([np.ones_like(x), x, y, z, x * x, x * y, x * z, y * y, y * z, z * z], 1)
out["mono_d0"] = full[:, :1]
out["mono_d1"] = full[:, :4]
out["mono_d2"] = full[:, :10]

# direct evaluation of a tiny bh3 interpolant with a degree-1 polynomial
src = rng.uniform(0, 1, size=(40, 3))
w = rng.normal(size=40)
c = rng.normal(size=4)
trg = rng.uniform(-0.2, 1.2, size=(25, 3))
R = np.linalg.norm(trg[:, None, :] - src[None, :, :], axis=2)
out["de_src"], out["de_w"], out["de_c"], out["de_trg"] = src, w, c, trg
out["de_bh3_y"] = (-R) @ w + c[0] + trg @ c[1:]
Ra = np.linalg.norm((trg @ aniso.T)[:, None, :] - (src @ aniso.T)[None, :, :], axis=2)
out["de_th3_aniso_y"] = (Ra**3) @ w + c[0] + trg @ c[1:]
out["de_exp_y"] = (1.0 * np.exp(-3.0 * R / 2.0)) @ w + c[0]   # degree 0: only c[0]

# direct operator [[A + eta I, P], [P^T, 0]] x for bh3 degree 1, nugget 0.01
n = 30
p = rng.uniform(0, 1, size=(n, 3))
xw = rng.normal(size=n + 4)
Rp = np.linalg.norm(p[:, None, :] - p[None, :, :], axis=2)
A = -Rp + 0.01 * np.eye(n)
P = np.concatenate([np.ones((n, 1)), p], 1)
out["op_pts"], out["op_x"] = p, xw
out["op_y"] = np.concatenate([A @ xw[:n] + P @ xw[n:], P.T @ xw[:n]])

# box: bbox corners through A, 1.01 x max width (src/fmm/utility.hpp:19-33)
bmin = np.array([-0.3, 0.1, 0.2])
bmax = np.array([1.1, 0.4, 2.0])
corners = np.array([[bmax[0] if i & 1 else bmin[0], bmax[1] if i & 2 else bmin[1],
                     bmax[2] if i & 4 else bmin[2]] for i in range(8)]) @ aniso.T
lo, hi = corners.min(0), corners.max(0)
out["box_min"], out["box_max"] = bmin, bmax
out["box_center"] = lo + (hi - lo) / 2
out["box_width"] = np.array([1.01 * (hi - lo).max()])

np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "closed_forms.npz"), **out)
print("wrote", len(out), "arrays")
