"""Write golden vectors for the oracle functions that no closed form or
cross-path equality pins (reference repository's accusim, imported read-only
in the build container only; the fixtures are data: inputs and outputs).

Covered (one .npz each, small):
  plummer_n64_seed1.npz       examples/star_cluster.py:14-37 sampler outputs
  bh_theta05_n300.npz         tree.py:45-78 forces/energy at theta=0.5 and 0.3
  argon864_vv20.npz           20 velocity-Verlet steps of the 864-atom argon
                              case (per-step PE, KE, momentum; final state)
  richardson_argon108.npz     accuracy.py:191-221 on a 108-atom argon run

Usage: python tests/golden/make_golden.py /path/to/reference
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    sys.path.insert(0, ref)
    sys.path.insert(0, os.path.join(ref, "examples"))
    import accusim
    from accusim import BarnesHut, Gravity, LennardJones, Simulation, System, VelocityVerlet
    from accusim.accuracy import timestep_convergence
    from star_cluster import plummer

    s, a, M = plummer(n=64, seed=1)
    np.savez(os.path.join(HERE, "plummer_n64_seed1.npz"), pos=s.pos, vel=s.vel, mass=s.masses, a=a, M=M)

    s, a, M = plummer(n=300, seed=1)
    out = {"pos": s.pos, "mass": s.masses, "softening": 0.01 * a}
    for th in (0.5, 0.3):
        F, U = BarnesHut(softening=0.01 * a, theta=th).compute(s)
        out[f"F_{th}"] = F
        out[f"U_{th}"] = U
    F0, U0 = Gravity(softening=0.01 * a).compute(s)
    out["F_direct"], out["U_direct"] = F0, U0
    np.savez(os.path.join(HERE, "bh_theta05_n300.npz"), **out)

    def argon(n_cells, seed):
        m = 39.948 * 1.66053906660e-27
        L = (4 * n_cells ** 3 * m / 1374.0) ** (1 / 3)
        base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) + 0.25
        pts = np.array([b + [i, j, k] for i in range(n_cells) for j in range(n_cells) for k in range(n_cells)
                        for b in base])
        s = System(3, box=L, periodic=True)
        s.add_species("Ar", m)
        s.add("Ar", pts * L / n_cells)
        s.maxwell_boltzmann(94.4, np.random.default_rng(seed))
        return s

    params = {"Ar": (119.8 * 1.380649e-23, 3.405e-10)}
    s = argon(6, 64)
    pos0, vel0 = s.pos.copy(), s.vel.copy()
    lj = LennardJones(params, cutoff=2.5, skin=0.3)
    sim = Simulation(s, [lj], integrator=VelocityVerlet(), dt=10e-15)
    F0 = sim.force.copy()
    tr = sim.run(steps=20, positions=False)
    np.savez(os.path.join(HERE, "argon864_vv20.npz"), pos0=pos0, vel0=vel0, F0=F0, U0=tr.potential[0],
             pe=np.array(tr.potential), ke=np.array(tr.kinetic), mom=np.array(tr.momentum),
             pos=s.pos, vel=s.vel, F=sim.force, rebuilds=lj.rebuilds, box=s.box)

    s = argon(3, 5)
    sim = Simulation(s, [LennardJones(params, cutoff=2.5, pairlist_threshold=10 ** 9)],
                     integrator=VelocityVerlet(), dt=10e-15)
    pos0, vel0 = s.pos.copy(), s.vel.copy()
    sim.run(steps=5, positions=False)
    chk = timestep_convergence(sim, 50 * 10e-15)
    np.savez(os.path.join(HERE, "richardson_argon108.npz"), pos0=pos0, vel0=vel0, box=s.box,
             status=chk.status, value=chk.value, detail=chk.detail)
    print("wrote golden vectors with accusim", getattr(accusim, "__version__", "?"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
