"""NumPy oracle of the learned DSF -> Ewald pair correction (test infrastructure only).

The correction has no counterpart in the reference (it is new in this build); its
parity is pinned by (a) a finite-difference check of the analytic gradient below,
(b) exact antisymmetry (sum of correction forces = 0) and (c) agreement of the GPU
forward/backward kernels with these functions.

Model (per directed pair i -> j within rc, d = r_j - r_i, r = |d|):
  n_i   = sum_{j != i} sw(r_ij),  sw = 1 below r_on, cosine to 0 at r_off
  x     = [r/rc, q_i q_j / e^2, (n_i + n_j)/n_scale, |n_i - n_j|/n_scale, (q_i + q_j)/e]
          (symmetric in i, j)
  h1    = silu(W1 x + b1)            (64)
  h2    = silu(W2 h1 + b2)           (64; W2 and h1 rounded to bfloat16 on the GPU)
  s     = (w3 . h2 + b3) * (1 - (r/rc)^2)^2
  F_i  += F0 * s * d / r
Loss per snapshot: L = mean over atoms and components of ((F_corr - dF_target) / F0)^2.
"""
from __future__ import annotations

import numpy as np

H = 64
N_IN = 5


def bf16(x):
    """Round float64/float32 values to bfloat16 (round to nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    sg = 1.0 / (1.0 + np.exp(-z))
    return sg * (1.0 + z * (1.0 - sg))


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    return {
        "W1": rng.normal(0, 1 / np.sqrt(N_IN), (H, N_IN)),
        "b1": np.zeros(H),
        "W2": rng.normal(0, 1 / np.sqrt(H), (H, H)),
        "b2": np.zeros(H),
        "w3": rng.normal(0, 0.1 / np.sqrt(H), H),
        "b3": np.zeros(1),
    }


def switch(r, r_on, r_off):
    x = np.clip((r - r_on) / (r_off - r_on), 0.0, 1.0)
    return 0.5 * (1.0 + np.cos(np.pi * x))


def directed_pairs(pos, box, rc):
    """All ordered pairs (i, j), i != j, with minimum-image r < rc (brute force)."""
    d = pos[None, :, :] - pos[:, None, :]
    d -= box * np.round(d / box)
    r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
    np.fill_diagonal(r, np.inf)
    i, j = np.nonzero(r < rc)
    return i, j, d[i, j], r[i, j]


def coordination(pos, box, r_on, r_off):
    d = pos[None, :, :] - pos[:, None, :]
    d -= box * np.round(d / box)
    r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
    np.fill_diagonal(r, np.inf)
    return switch(r, r_on, r_off).sum(1)


def features(i, j, r, q, n, rc, e2, n_scale):
    return np.stack([r / rc, q[i] * q[j] / e2, (n[i] + n[j]) / n_scale, np.abs(n[i] - n[j]) / n_scale,
                     (q[i] + q[j]) / np.sqrt(e2)], 1)


def forward(p, x, r, rc, emulate_bf16=True):
    """Per-pair s (with the envelope) and the cached activations."""
    z1 = x @ p["W1"].T + p["b1"]
    h1 = silu(z1)
    W2 = bf16(p["W2"]) if emulate_bf16 else p["W2"]
    h1q = bf16(h1) if emulate_bf16 else h1
    z2 = h1q @ W2.T + p["b2"]
    h2 = silu(z2)
    raw = h2 @ p["w3"] + p["b3"][0]
    env = (1.0 - (r / rc) ** 2) ** 2
    return raw * env, (z1, h1, h1q, W2, z2, h2, raw, env)


def correction_forces(p, n_atoms, i, d, r, x, rc, F0, emulate_bf16=True):
    s, cache = forward(p, x, r, rc, emulate_bf16)
    F = np.zeros((n_atoms, 3))
    np.add.at(F, i, F0 * (s / r)[:, None] * d)
    return F, s, cache


def loss_and_grad(p, n_atoms, i, d, r, x, rc, F0, target, emulate_bf16=True):
    """Loss and its gradient for one snapshot (target = F_Ewald - F_DSF, in newtons)."""
    F, s, (z1, h1, h1q, W2, z2, h2, raw, env) = correction_forces(p, n_atoms, i, d, r, x, rc, F0, emulate_bf16)
    e = (F - target) / F0
    m = e.size
    L = float(np.sum(e ** 2) / m)
    g_atom = 2.0 * e / m                      # dL/d(F_i/F0)
    ds = np.einsum("pk,pk->p", g_atom[i], d) / r   # dL/ds (s already includes the envelope)
    draw = ds * env
    g = {"w3": h2.T @ draw, "b3": np.array([draw.sum()])}
    dz2 = np.outer(draw, p["w3"]) * dsilu(z2)
    g["b2"] = dz2.sum(0)
    # emulate_bf16: the GPU's backward GEMMs take bfloat16 operands (dz2, dz1 and the
    # features x), rounded at the same points here
    dz2q = bf16(dz2) if emulate_bf16 else dz2
    g["W2"] = dz2q.T @ h1q
    dh1 = dz2q @ W2
    dz1 = dh1 * dsilu(z1)
    g["b1"] = (bf16(dz1) if emulate_bf16 else dz1).sum(0)
    if emulate_bf16:
        dz1, x = bf16(dz1), bf16(x)
    g["W1"] = dz1.T @ x
    return L, g, F
