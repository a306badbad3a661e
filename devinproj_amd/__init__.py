"""MI355X (gfx950) implementation of accusim's non-bonded force path.

Drop-in classes with the reference's Force API (accusim/forces.py, tree.py),
a device-resident fused MD engine, and the accuracy-report rows on this path.
The HIP library (_accuhip.so) is required; there is no CPU fallback.
"""
from .accuracy import AccuracyReport, Check, timestep_convergence
from .batch import BrickEngine
from .forces import (DampedShiftedCoulomb, Ewald, Force, Gravity, LennardJones, NeighbourList, VerletList,
                     cell_list_pairs, kvector_columns)
from .integrators import Brownian, Integrator, Langevin, NoseHoover, Rattle, VelocityVerlet, Yoshida4
from . import autotune, batch, workloads
from .learned import LearnedDSF, PairCorrection, train_correction
from .nbody import NBodyEngine
from .simulation import Simulation, Trajectory
from .system import Species, System
from .tree import BarnesHut, direct_subset
from .units import G, k_B, k_e, si

__all__ = [
    "AccuracyReport", "BarnesHut", "BrickEngine", "Brownian", "Check", "DampedShiftedCoulomb", "Ewald", "Force", "G",
    "Gravity", "Integrator", "Langevin", "LearnedDSF", "LennardJones", "NBodyEngine", "NeighbourList", "NoseHoover",
    "PairCorrection", "Rattle", "Simulation", "Species", "System",
    "Trajectory", "VelocityVerlet", "train_correction",
    "VerletList", "Yoshida4", "cell_list_pairs", "direct_subset", "k_B", "k_e", "kvector_columns", "si",
    "timestep_convergence",
]
