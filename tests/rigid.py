"""Rigid-molecule fluids for the Rattle tests (reduced units: mass 1, LJ eps 1, sigma 0.8).

shape "dimer": 2 atoms, bond 1; "triangle": 3 atoms, all three sides 1 (a 3 x 3 cluster
with a closed ring); "chain": 4 atoms in a planar zigzag with bonds 1 (a 3-bond path,
flexible angles); "chain8": the same zigzag of 8 atoms (7 bonds). Molecule centres sit on a cubic grid of the periodic box, positions are
wrapped (bonds straddle the faces), velocities are a random rigid motion of each molecule
(translation plus rotation about its centre), so they satisfy the constraints."""
import numpy as np

from oracle import accusim_np as A

SHAPES = {
    "dimer": (np.array([[-0.5, 0, 0], [0.5, 0, 0]]), [(0, 1)]),
    "triangle": (np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]]), [(0, 1), (1, 2), (0, 2)]),
    "chain": (np.array([[-1.2, -0.3, 0], [-0.4, 0.3, 0], [0.4, -0.3, 0], [1.2, 0.3, 0]]), [(0, 1), (1, 2), (2, 3)]),
    "chain8": (np.array([[0.8 * i - 2.8, 0.3 * (-1) ** (i + 1), 0] for i in range(8)]), [(i, i + 1) for i in range(7)]),
}


def random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.diag(r))


def fluid(shape, n_mol, rho, seed=0, v_scale=0.5):
    """(System, constraints [(id_a, id_b, length)]) of n_mol molecules at number density rho."""
    body, bonds = SHAPES[shape]
    body = body - body.mean(0)
    rng = np.random.default_rng(seed)
    nat = len(body)
    L = (nat * n_mol / rho) ** (1 / 3)
    m = int(np.ceil(n_mol ** (1 / 3)))
    g = (np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n_mol] + 0.5) * (L / m)
    s = A.System(3, box=L, periodic=True)
    s.add_species("A", 1.0)
    ids = np.empty((n_mol, nat), dtype=int)
    pos = np.empty((n_mol, nat, 3))
    vel = np.empty((n_mol, nat, 3))
    for k in range(n_mol):
        x = body @ random_rotation(rng).T
        w = rng.normal(0, v_scale, 3)
        pos[k] = (g[k] + x) % L
        vel[k] = rng.normal(0, v_scale, 3) + np.cross(w, x)
    for j in range(nat):  # atom j of every molecule, so ids of one molecule are n_mol apart
        ids[:, j] = s.add("A", pos[:, j], vel=vel[:, j])
    s.vel -= s.vel.mean(0)
    cons = [(int(ids[k, a]), int(ids[k, b]), float(np.linalg.norm(body[a] - body[b])))
            for k in range(n_mol) for a, b in bonds]
    return s, cons
