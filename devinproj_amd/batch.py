"""Brick-tiled, LDS-staged MD engine for one or many periodic systems.

`BrickEngine([systems], [forces per system], integrator, dts)` advances several
independent systems in one device arena (the C ABI `accu_bmd_*`). Typical uses:
  - one large system (fast path for the 1M-atom LJ workload);
  - the dt, dt/2, dt/4 re-runs of accusim's timestep_convergence
    (accuracy.py:191-221) for many systems at once: `richardson()` orders the
    trajectories so that the ones advanced at each tick form a prefix, runs them in
    one schedule (4 ticks per coarse step: dt/4 runs every tick, dt/2 every other
    tick, dt once), optionally as replays of one captured hipGraph, and evaluates
    the reference's Richardson formula on the final positions.
Brick shapes (the LDS tile) are per system; `autotune.choose_tile` picks them (and the workgroup size) by
timing candidates.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from .accuracy import richardson_check
from .forces import DampedShiftedCoulomb, LennardJones, _erfc
from .integrators import Brownian, Langevin, NoseHoover, Rattle, VelocityVerlet, Yoshida4, constraint_clusters
from .simulation import Trajectory

LOG_W = 12  # PE, KE, Px, Py, Pz, rebuilds, Lx, Ly, Lz, 0, 0, 0


def default_bricks(s, rc_list, cap_atoms=1500):
    """Largest brick (cells per axis <= 3) whose staged region stays near cap_atoms."""
    nc = np.floor(np.asarray(s.box, float) / rc_list).astype(int)
    per_cell = s.n / float(np.prod(nc))
    best = (1, 1, 1)
    for b in [(1, 1, 1), (1, 1, 2), (1, 2, 2), (2, 2, 2), (2, 2, 3), (2, 3, 3), (3, 3, 3)]:
        if any(b[d] > nc[d] for d in range(3)):
            continue
        region = np.prod(np.asarray(b) + 2) * per_cell
        if region <= cap_atoms:
            best = b
    return best


class BrickEngine:
    """Device-resident MD of one or many periodic systems (csrc/brick.hip).

    `slack` is the relative headroom of the neighbour-list and LDS region capacities
    over what the first build measures. A region or list that outgrows it later is
    reported as an error at fetch(). The LDS region sets how many workgroups share a CU,
    so a larger slack costs occupancy.

    Every record (the set-up state and then every `record_every` steps of `run`) holds
    each system's potential and kinetic energy, momentum and angular momentum, summed in
    float64 on the device. With `snapshots` > 0, a device ring of that many frames also
    keeps the positions and velocities of every `snapshot_every`-th record (the set-up
    state is frame 0). `trajectory(k)` turns both into the reference's Trajectory
    (simulation.py:14-39), which `devinproj_amd.io.save` writes in its .npz layout.
    """

    def __init__(self, systems, forces, integrator=None, dts=None, precision="fp64", bricks=None, slack=0.25,
                 log_cap=20_000, threads=256, snapshots=0, snapshot_every=1):
        self.systems = list(systems)
        S = len(self.systems)
        if len(forces) != S:
            raise ValueError("one force list per system")
        if not isinstance(dts, (list, tuple, np.ndarray)):
            dts = [dts] * S
        self.dts = [float(d) for d in dts]
        self.integrator = integrator or VelocityVerlet()
        ig = self.integrator
        # integrator mode of the device engine (accu_bmd_set_integrator)
        if isinstance(ig, Yoshida4):
            w = Yoshida4.weights()
            self.ticks = (3, [S, S, S], [w[0], w[1], w[2]], [w[1], w[2], w[0]])
            self._mode = (0, 0.0, 0.0, 0.0, 0)
        elif isinstance(ig, VelocityVerlet):
            self.ticks = (1, [S], [1.0], [1.0])
            self._mode = (0, 0.0, 0.0, 0.0, 0)
        elif isinstance(ig, (Langevin, NoseHoover, Brownian, Rattle)):
            self.ticks = (1, [S], [1.0], [1.0])
            if isinstance(ig, Rattle):
                self._mode = (4, 0.0, 0.0, 0.0, 0)
            elif isinstance(ig, Langevin):
                self._mode = (1, ig.T, ig.gamma, 0.0, ig.seed)
            elif isinstance(ig, Brownian):
                self._mode = (2, ig.T, 0.0, 0.0, ig.seed)
            else:
                if ig.Q is not None:
                    raise ValueError("the device Nose-Hoover starts from a fresh thermostat (xi = eta = 0)")
                self._mode = (3, ig.T, 0.0, ig.tau, 0)
        else:
            raise ValueError("the brick engine runs VelocityVerlet, Yoshida4, Langevin, NoseHoover, Brownian or Rattle")
        self.precision = precision
        L.lib()
        for s in self.systems:
            if s.dim != 3 or not s.periodic:
                raise ValueError("the brick engine needs 3D periodic boxes")
        natoms = np.array([s.n for s in self.systems], np.int32)
        pos = np.ascontiguousarray(np.concatenate([s.pos for s in self.systems]), float)
        vel = np.ascontiguousarray(np.concatenate([s.vel for s in self.systems]), float)
        sp = np.ascontiguousarray(np.concatenate([s.sp for s in self.systems]), np.int32)
        box = np.ascontiguousarray(np.stack([np.asarray(s.box, float) for s in self.systems]))
        h = L.C.c_void_p()
        L.call("accu_bmd_create", L.C.byref(h), S, L.ptr(natoms), L.ptr(pos), L.ptr(vel), L.ptr(sp), L.ptr(box),
               L.PRECISIONS[precision])
        self._h = h
        self.bricks = []
        for k, (s, fl) in enumerate(zip(self.systems, forces)):
            lj = [f for f in fl if isinstance(f, LennardJones)]
            dsf = [f for f in fl if isinstance(f, DampedShiftedCoulomb)]
            if len(lj) > 1 or len(dsf) > 1 or len(lj) + len(dsf) != len(fl) or not fl:
                raise ValueError("each system takes one LennardJones and/or one DampedShiftedCoulomb")
            radii, skins = [], []
            mass = np.ascontiguousarray([x.mass for x in s.species], float)
            charge = np.ascontiguousarray([x.charge for x in s.species], float)
            if lj:
                rc, sk = lj[0].list_radius(s)
                radii.append(rc)
                skins.append(sk)
            if dsf:
                if dsf[0].exclude:
                    raise ValueError("exclusions are not supported by the device engines")
                radii.append(dsf[0].rcut)
                skins.append(dsf[0].skin)
            skin = min(skins)
            rc_list = max(r + skin for r in radii)
            b = tuple(bricks[k]) if bricks is not None and bricks[k] is not None else default_bricks(s, rc_list)
            self.bricks.append(b)
            bd = np.ascontiguousarray(b, np.int32)
            L.call("accu_bmd_set_system", h, k, len(s.species), L.ptr(mass), L.ptr(charge), self.dts[k],
                   float(rc_list), float(skin), L.ptr(bd))
            if lj:
                eps, sig2, rc2, shift = lj[0].device_tables(s)
                L.call("accu_bmd_set_lj", h, k, len(s.species), L.ptr(eps), L.ptr(sig2), L.ptr(rc2), shift)
            if dsf:
                f = dsf[0]
                erc = float(_erfc(np.array(f.alpha * f.rcut)))
                L.call("accu_bmd_set_dsf", h, k, float(f.alpha), float(f.rcut), erc, float(f.k))
            if self._mode[0] == 2:
                D = np.ascontiguousarray([x.diffusion for x in s.species], float)
                L.call("accu_bmd_set_diffusion", h, k, len(s.species), L.ptr(D))
        if self._mode[0] == 4:
            self._set_constraints(ig)
        mode, T, gamma, tau, seed = self._mode
        L.call("accu_bmd_set_integrator", h, mode, float(T), float(gamma), float(tau),
               L.C.c_ulonglong(int(seed) & 0xFFFFFFFFFFFFFFFF))
        self.log_cap = log_cap
        self.threads = threads
        L.call("accu_bmd_set_threads", h, int(threads))
        self.snapshots, self.snapshot_every = int(snapshots), int(snapshot_every)
        if self.snapshots:
            L.call("accu_bmd_set_snapshots", h, self.snapshots, self.snapshot_every)
        L.call("accu_bmd_init", h, float(slack), int(log_cap))
        self.steps = 0
        # simulated time of every record, per system (the set-up record is at t = 0)
        self._clock = np.zeros(S)
        self._rec_times = [self._clock.copy()]

    def _set_constraints(self, rat):
        """Rattle's bonds as connected clusters per system (accu_bmd_set_constraints): the
        bonds whose two atoms are in a system apply to it (ids are per system, as in the
        reference's Rattle.step), each system counts them as constrained degrees of freedom."""
        coff, aoff, la, lb, ln, uatom, csys = [0], [0], [], [], [], [], []
        off = 0
        for k, s in enumerate(self.systems):
            where = {int(pid): n for n, pid in enumerate(s.ids)}
            cons = [(where[a], where[b], d) for a, b, d in rat.constraints if a in where and b in where]
            s.constrained_dof = len(cons)
            ia = [c[0] for c in cons]
            ib = [c[1] for c in cons]
            for g in constraint_clusters(ia, ib):
                atoms = []
                loc = {}
                for q in g:
                    for x in (ia[q], ib[q]):
                        if x not in loc:
                            loc[x] = len(atoms)
                            atoms.append(x)
                    la.append(loc[ia[q]])
                    lb.append(loc[ib[q]])
                    ln.append(cons[q][2])
                uatom.extend(off + x for x in atoms)
                coff.append(len(la))
                aoff.append(len(uatom))
                csys.append(k)
            off += s.n
        if not la:
            raise ValueError("Rattle: no constraint has both atoms in a system")
        arr = [np.ascontiguousarray(x, np.int32) for x in (coff, aoff, la, lb)]
        lens = np.ascontiguousarray(ln, float)
        ua, cs = np.ascontiguousarray(uatom, np.int32), np.ascontiguousarray(csys, np.int32)
        L.call("accu_bmd_set_constraints", self._h, len(csys), *[L.ptr(a) for a in arr], L.ptr(lens), L.ptr(ua),
               L.ptr(cs), float(rat.tol), int(rat.max_iter))

    def constraint_state(self):
        """(worst relative bond-length error after any SHAKE so far, error code, cluster,
        residual at the failure) of a Rattle engine; raises the reference's RuntimeError
        when a SHAKE or RATTLE solve failed."""
        out = np.zeros(2)
        err = np.zeros(2, np.int32)
        L.call("accu_bmd_get_constraints", self._h, L.ptr(out), L.ptr(err))
        code = int(err[0])
        # the failing cluster, indexed as _set_constraints listed them (integrators.
        # constraint_clusters order, system after system); None while nothing failed
        self.failed_cluster = int(err[1]) if code else None
        if code == 1:
            raise RuntimeError("SHAKE did not converge: reduce dt")
        if code == 2:
            raise RuntimeError(f"SHAKE stalled at relative residual {out[1]:.1e} > tol {self.integrator.tol:.0e}: "
                               "raise tol (float64 limit ~ eps * coordinate size / bond length) or reduce dt")
        if code == 3:
            raise RuntimeError(f"RATTLE: singular velocity system in constraint cluster {int(err[1])}")
        return float(out[0])

    def run(self, steps, record_every=1, use_graph=False, sync=True, schedule=None):
        nt, active, wc, wn = schedule or self.ticks
        a = np.ascontiguousarray(active, np.int32)
        c = np.ascontiguousarray(wc, float)
        n = np.ascontiguousarray(wn, float)
        L.call("accu_bmd_run", self._h, int(steps), int(nt), L.ptr(a), L.ptr(c), L.ptr(n), int(record_every),
               int(bool(use_graph)))
        # time per coarse step of system k: its dt times the drift weights of the ticks it takes part in
        S = len(self.systems)
        per = np.array([self.dts[k] * sum(n[t] for t in range(nt) if k < a[t]) for k in range(S)])
        if record_every > 0:
            for c_ in range(1, steps + 1):
                if c_ % record_every == 0 or c_ == steps:
                    self._rec_times.append(self._clock + c_ * per)
        self._clock = self._clock + steps * per
        self.steps += steps
        if sync:
            self.sync()
            if isinstance(self.integrator, Rattle):
                # a failed SHAKE or RATTLE raises here, at the end of the run that hit it
                # (the device keeps stepping the other clusters until the run ends)
                self.constraint_state()
        return self

    def sync(self):
        L.call("accu_bmd_sync", self._h)

    def fetch(self, state=True):
        S = len(self.systems)
        N = sum(s.n for s in self.systems)
        pos = np.empty((N, 3)) if state else None
        vel = np.empty((N, 3)) if state else None
        F = np.empty((N, 3)) if state else None
        log = np.zeros((S, self.log_cap, LOG_W))
        stats = np.zeros((S, 4), np.int64)
        info = np.zeros(16, np.int64)
        L.call("accu_bmd_get", self._h, L.ptr(pos), L.ptr(vel), L.ptr(F), L.ptr(log), self.log_cap, L.ptr(stats),
               L.ptr(info))
        khz = max(int(info[12]), 1)
        ends = [1e3 * int(info[5 + k]) / khz for k in (0, 1, 2, 4)]  # phase ends after the rebuild's start
        names = ("binning", "scan", "grouping", "order_and_gather")  # grouping: 0 unless a bucket overflowed
        self.info = {"max_neighbours": int(info[0]), "capacity": int(info[1]), "region_capacity": int(info[2]),
                     "list_entries": int(info[4]), "bucket_overflows": int(info[10]), "bucket_cap": int(info[11]),
                     "rebuild_phase_us": {n: round(e - (ends[k - 1] if k else 0.0), 1)
                                          for k, (n, e) in enumerate(zip(names, ends))}}
        self.rebuilds = stats[:, 0].copy()
        self.logs = [log[k, :int(stats[k, 1])].copy() for k in range(S)]
        if isinstance(self.integrator, Rattle):
            self.integrator.max_violation = max(self.integrator.max_violation, self.constraint_state())
        if state:
            off = 0
            self.forces_out = []
            for s in self.systems:
                s.pos[...] = pos[off:off + s.n]
                s.vel[...] = vel[off:off + s.n]
                self.forces_out.append(F[off:off + s.n])
                off += s.n
        return self.logs

    def snapshot_frames(self):
        """(positions [F][N][3], velocities [F][N][3], record index [F]) of the stored
        frames of the whole arena, in user order; frames beyond the ring are dropped."""
        taken = L.C.c_int(0)
        L.call("accu_bmd_get_snapshots", self._h, 0, 0, None, None, None, L.C.byref(taken))
        nf = min(taken.value, self.snapshots)
        N = sum(s.n for s in self.systems)
        pos = np.empty((nf, N, 3))
        vel = np.empty((nf, N, 3))
        rec = np.empty(nf, np.int32)
        if nf:
            L.call("accu_bmd_get_snapshots", self._h, 0, nf, L.ptr(pos), L.ptr(vel), L.ptr(rec), L.C.byref(taken))
        return pos, vel, rec

    def trajectory(self, k=0):
        """System k's records (and stored frames) as a Trajectory, after fetch().

        Per record: t, kinetic, potential, momentum, angular_momentum from the device
        log; internal, released and bath are 0 and mass, charge and counts constant
        (no reactions on this path, simulation.py:76-93). Frames fill positions,
        velocities and species; `frame_records` gives each frame's record index."""
        s = self.systems[k]
        lg = self.logs[k]
        n = len(lg)
        tr = Trajectory()
        tr.t = [float(self._rec_times[r][k]) for r in range(n)]
        tr.kinetic = lg[:, 1].tolist()
        tr.potential = lg[:, 0].tolist()
        tr.internal = [0.0] * n
        tr.released = [0.0] * n
        tr.momentum = [lg[r, 2:5].copy() for r in range(n)]
        tr.angular_momentum = [lg[r, 6:9].copy() for r in range(n)]
        tr.mass = [float(np.sum(s.masses))] * n
        tr.charge = [float(np.sum(s.charges))] * n
        tr.counts = [s.counts() for _ in range(n)]
        tr.bath = lg[:, 9].tolist()  # Nose-Hoover bath energy (0 for the other integrators)
        tr.frame_records = []
        if self.snapshots:
            off = sum(x.n for x in self.systems[:k])
            pos, vel, rec = self.snapshot_frames()
            for f in range(len(rec)):
                if 0 <= rec[f] < n:
                    tr.positions.append(pos[f, off:off + s.n].copy())
                    tr.velocities.append(vel[f, off:off + s.n].copy())
                    tr.species.append(s.sp.copy())
                    tr.frame_records.append(int(rec[f]))
        return tr

    def thermostat_state(self):
        """Nose-Hoover state per system: rows (xi, eta, Q, the velocity scales of the last
        two half steps, k2)."""
        out = np.zeros((len(self.systems), 6))
        L.call("accu_bmd_get_thermostat", self._h, L.ptr(out))
        return out

    def energy_drift(self, k=0):
        lg = self.logs[k]
        E = lg[:, 0] + lg[:, 1]
        scale = max(abs(E[0]), np.max(np.abs(lg[:, 1])), np.max(np.abs(lg[:, 0])), 1e-300)
        return float(np.max(np.abs(E - E[0])) / scale)

    def set_timing(self, enable=True):
        L.call("accu_bmd_set_timing", self._h, int(bool(enable)))

    def timing(self):
        ms, k = L.C.c_double(0.0), L.C.c_longlong(0)
        L.call("accu_bmd_timing", self._h, L.C.byref(ms), L.C.byref(k))
        return ms.value, k.value

    def close(self):
        h = getattr(self, "_h", None)
        lib = getattr(L, "_lib", None) if L is not None else None  # None at interpreter exit
        if h is not None and lib is not None:
            lib.accu_bmd_destroy(h)
        self._h = None

    def __del__(self):
        self.close()


def richardson(systems, make_forces, dt, duration_steps, precision="mixed", bricks=None, use_graph=True, block=10,
               log=None, threads=256):
    """timestep_convergence (accuracy.py:191-221) for many systems in one batched run.

    Every system is re-run from its current state at dt, dt/2 and dt/4 for
    `duration_steps` coarse steps (= duration_steps * k steps at dt/k); returns one
    Check per system and the engine (for timing)."""
    S = len(systems)
    # order: all dt/4 replicas, then dt/2, then dt, so each tick advances a prefix
    order = [(k, 4) for k in range(S)] + [(k, 2) for k in range(S)] + [(k, 1) for k in range(S)]
    reps = [systems[k].copy() for k, _ in order]
    dts = [dt / m for _, m in order]
    forces = [make_forces(r) for r in reps]
    br = None if bricks is None else [bricks[k] for k, _ in order]
    eng = BrickEngine(reps, forces, VelocityVerlet(), dts, precision=precision, bricks=br, threads=threads)
    sched = (4, [S, S, 2 * S, 3 * S], [1.0] * 4, [1.0] * 4)
    left = duration_steps
    t0 = time.perf_counter()
    while left > 0:
        blk = min(block, left) if use_graph else left
        eng.run(blk, record_every=10 ** 9, use_graph=use_graph and blk == block, sync=False, schedule=sched)
        left -= blk
    eng.sync()
    eng.run_seconds = time.perf_counter() - t0
    eng.fetch(state=True)
    checks = []
    for k in range(S):
        finals = [reps[k + j * S].pos.copy() for j in (2, 1, 0)]  # dt, dt/2, dt/4
        checks.append(richardson_check(finals, duration_steps, 2))
    return checks, eng, reps
