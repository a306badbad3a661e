"""Golden accuracy-report rows from the reference's own AccuracyReport.build
(accusim/accuracy.py:59-189) and advisor.timescales (advisor.py:9-48), for the small runs of
tests/report_cases.py. The reference is imported read-only in the build container only;
the fixture holds data: each case's initial state and the rows (name, status, value).

Usage: python tests/golden/make_report_golden.py /path/to/reference
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import report_cases as RC  # noqa: E402


def main(ref):
    sys.path.insert(0, ref)
    import accusim
    from accusim import advisor, integrators, tree
    from accusim.accuracy import AccuracyReport
    M = type("Ref", (), {})
    for name in ("System", "Simulation", "LennardJones", "DampedShiftedCoulomb"):
        setattr(M, name, getattr(accusim, name, None) or getattr(accusim.forces, name))
    for name in ("VelocityVerlet", "NoseHoover", "Rattle"):
        setattr(M, name, getattr(integrators, name))
    out = {}
    for case in RC.CASES:
        st = RC.state(case)
        sim, steps = RC.build(M, case, st, tree=tree)
        sim.run(steps=steps, positions=False)
        rep = AccuracyReport.build(sim)
        ts = advisor.timescales(sim)
        names = [c.name for c in rep.checks]
        out[f"{case}_names"] = np.array(names)
        out[f"{case}_status"] = np.array([c.status for c in rep.checks])
        out[f"{case}_value"] = np.array([np.nan if c.value is None else c.value for c in rep.checks])
        out[f"{case}_ts_kinds"] = np.array(sorted(ts))
        out[f"{case}_ts_values"] = np.array([ts[k] for k in sorted(ts)])
        out[f"{case}_final_pos"] = sim.system.pos.copy()
        print(case, list(zip(names, [c.status for c in rep.checks])), ts)
    np.savez(os.path.join(HERE, "report_rows.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
