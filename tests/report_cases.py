"""Small runs whose accuracy reports are compared row by row (accuracy.py:59-189).

Each case is an initial state (arrays built from the oracle's seeded builders), a force
list, an integrator, dt and a step count. `build(M, case)` makes the Simulation from any
module M with the reference's class names: the reference's `accusim` (in the container
only, to write the golden rows), the oracle, or `devinproj_amd` (GPU forces).

  argon        108 LJ argon atoms, velocity Verlet at 10 fs, 30 steps
  plummer      300-star Plummer sphere, Barnes-Hut theta 0.5, velocity Verlet, 5 steps
  nacl         216 ions, LJ + DSF (rc 9 A, alpha 0.2 / A), velocity Verlet at 1 fs, 5 steps
  nosehoover   108 argon atoms, Nose-Hoover at 120 K (period 0.2 ps), 40 steps
  rattle       32 rigid dimers (reduced units), Rattle tol 1e-10, 20 steps
"""
import numpy as np

from oracle import builders as B

CASES = ("argon", "plummer", "nacl", "nosehoover", "rattle")


def state(case):
    """Initial arrays of a case: pos, vel, sp, names, masses, charges, box (or None), extra."""
    extra = {}
    if case in ("argon", "nosehoover"):
        s = B.argon(3, seed=5, jitter=0.02)
    elif case == "plummer":
        s, a, M = B.plummer(n=300, seed=2)
        extra = {"a": a, "M": M}
    elif case == "nacl":
        s = B.rocksalt(6, 6, 6, seed=4)
        rng = np.random.default_rng(9)
        s.pos = (s.pos + rng.normal(0, 0.1e-10, s.pos.shape)) % s.box
    elif case == "rattle":
        import rigid
        s, cons = rigid.fluid("dimer", 32, 0.2, seed=4)
        extra = {"cons": np.array(cons, float)}
    else:
        raise KeyError(case)
    return {"pos": s.pos.copy(), "vel": s.vel.copy(), "sp": np.asarray(s.sp).copy(),
            "names": [x.name for x in s.species], "masses": np.array([x.mass for x in s.species]),
            "charges": np.array([x.charge for x in s.species]),
            "box": None if s.box is None else np.asarray(s.box, float).copy(), **extra}


def system(M, st):
    s = M.System(3, box=st["box"], periodic=st["box"] is not None)
    for name, m, q in zip(st["names"], st["masses"], st["charges"]):
        s.add_species(str(name), float(m), float(q))
    sp = np.asarray(st["sp"])
    start = 0
    while start < len(sp):  # contiguous runs of one species keep the particle order
        stop = start
        while stop < len(sp) and sp[stop] == sp[start]:
            stop += 1
        s.add(str(st["names"][sp[start]]), st["pos"][start:stop], vel=st["vel"][start:stop])
        start = stop
    return s


def build(M, case, st=None, tree=None, precision=None):
    """(Simulation, steps) of a case from module M; `tree` holds BarnesHut when M does not."""
    st = state(case) if st is None else st
    s = system(M, st)
    kw = {} if precision is None else {"precision": precision}
    if case in ("argon", "nosehoover"):
        forces = [M.LennardJones(B.argon_params(), cutoff=2.5, **kw)]
        dt, steps = 10e-15, 30
        integ = M.VelocityVerlet() if case == "argon" else M.NoseHoover(120.0, 0.2e-12)
        if case == "nosehoover":
            steps = 40
    elif case == "plummer":
        BH = getattr(tree or M, "BarnesHut")
        forces = [BH(softening=0.01 * float(st["a"]), theta=0.5)]
        t_cross = float(np.sqrt(float(st["a"]) ** 3 / (B.G_SI * float(st["M"]))))
        dt, steps, integ = 1e-3 * t_cross, 5, M.VelocityVerlet()
    elif case == "nacl":
        forces = [M.LennardJones(B.NACL_LJ, cutoff=2.5, **kw), M.DampedShiftedCoulomb(rcut=9e-10, alpha=0.2e10, **kw)]
        dt, steps, integ = 1e-15, 5, M.VelocityVerlet()
    elif case == "rattle":
        forces = [M.LennardJones({"A": (1.0, 0.8)}, cutoff=2.5, **kw)]
        cons = [(int(a), int(b), float(d)) for a, b, d in st["cons"]]
        dt, steps, integ = 1e-3, 20, M.Rattle(cons, tol=1e-10)
    sim = M.Simulation(s, forces, integrator=integ, dt=dt)
    return sim, steps
