"""python -m devinproj_amd model.json [--engine dropin|brick] [--precision fp64|mixed]
                                      [--snapshots N] [--save traj.npz] [--convergence]

Runs a JSON model (accusim/__main__.py, model.py) on the GPU force path and prints the
accuracy report. --engine dropin (default) drives the model's integrator through
Simulation.run with the device forces' compute(); --engine brick runs the whole step on
the device (periodic 3D boxes, LennardJones and/or DampedShiftedCoulomb). --save writes
the trajectory in the reference's .npz layout (io.py:13-43).
"""
from __future__ import annotations

import argparse
import sys

from . import model as M


def main(argv=None):
    ap = argparse.ArgumentParser(prog="devinproj_amd", description=__doc__.splitlines()[0])
    ap.add_argument("model")
    ap.add_argument("--engine", choices=("dropin", "brick"), default="dropin")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--snapshots", type=int, default=0, help="brick engine: frames kept on the device")
    ap.add_argument("--save", help="write the trajectory (.npz)")
    ap.add_argument("--convergence", action="store_true", help="estimate timestep error (3 extra runs, dropin)")
    a = ap.parse_args(argv)
    sim, model = M.load(a.model, precision=a.precision)
    if "duration" not in model:
        raise SystemExit("the model needs a duration")
    duration = M.si(model["duration"], "time", "duration")
    steps = int(round(duration / sim.dt))
    record_every = max(1, steps // 200)
    s = sim.system
    print(f"{s.n} particles, {len(s.species)} species, dt = {sim.dt:.3g} s, {steps} steps, engine {a.engine}")
    from .accuracy import AccuracyReport
    if a.engine == "dropin":
        tr = sim.run(steps=steps, record_every=record_every, positions=bool(a.save))
        report = AccuracyReport.build(sim, convergence_duration=duration if a.convergence else None)
    else:
        from .batch import BrickEngine
        eng = BrickEngine([s], [sim.forces], sim.integrator, sim.dt, precision=a.precision,
                          snapshots=a.snapshots, log_cap=steps // record_every + 16)
        eng.run(steps, record_every=record_every)
        eng.fetch()
        tr = eng.trajectory(0)
        report = AccuracyReport.from_run(tr, s, sim.forces, sim.integrator, sim.dt)
        eng.close()
    print(report)
    if a.save:
        from .io import save
        save(tr, s, a.save)
        print(f"trajectory written to {a.save}")
    return sim, report


if __name__ == "__main__":
    main(sys.argv[1:])
