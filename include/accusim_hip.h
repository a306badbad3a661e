/* C ABI of the MI355X (gfx950) force-path library for accusim.
 *
 * The reference (erezbz/DevinProj, package accusim) is pure Python/NumPy; its
 * interface for this path is the duck-typed Force.compute(system) -> (F, U)
 * (accusim/forces.py:1-21). The Python classes in devinproj_amd call the entry
 * points below through ctypes; each comment names the reference function whose
 * arithmetic the entry point replaces.
 *
 * Conventions
 *   - Pointers named pos/F/q/mass/... are DEVICE pointers (float64 unless noted),
 *     pos and F are [n][3] row-major like numpy arrays. Box/geometry arrays and
 *     anything named *_host are host pointers.
 *   - Return value 0 = ok; ACCU_EINVAL (-1) maps to ValueError, ACCU_ERUNTIME (-2)
 *     and ACCU_EOVERFLOW (-3) to RuntimeError; accu_last_error() has the message.
 *   - Work is enqueued on `stream` (a hipStream_t); outputs are caller-owned.
 *   - U arguments are device float64 scalars and may be NULL (no energy).
 *   - Handles are not thread-safe (one host thread at a time per handle). The entry points
 *     without a handle (pair forces, gravity, Ewald, accu_max_disp2) keep their scratch
 *     (reduction partials, grid-sum counter, Ewald buffers) per (device, stream): calls
 *     on different streams never share it, calls on one stream use it in stream order,
 *     and calls on one stream from several host threads at once must be serialised by
 *     the caller. The first call on a stream allocates that stream's scratch.
 */
#ifndef ACCUSIM_HIP_H
#define ACCUSIM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ACCU_ABI_VERSION 2

enum { ACCU_FP64 = 0, ACCU_FP32_PAIR = 1 };

typedef struct accu_nlist accu_nlist;

const char* accu_last_error(void);
int accu_version(void);
int accu_device_info(char* name, int len, int* cus, long long* hbm_bytes);
int accu_sync(void* stream);

/* ---- cell list / Verlet list: forces.py:279-321 (cell_list_pairs), 78-98 (VerletList) */
int accu_nlist_create(accu_nlist** out);
int accu_nlist_destroy(accu_nlist* h);
/* Bin atoms into ncell[0]*ncell[1]*ncell[2] cells of the region lo + [0, span)
 * (forces.py:282-292; the caller computes lo, span, ncell exactly as the reference
 * does), stable-sort them by cell, and store every pair with minimum-image
 * |d|^2 < rc^2 in both directions. box may be NULL when periodic == 0. */
int accu_nlist_build(accu_nlist* h, const double* pos, int n, const double* lo_host, const double* span_host,
                     const int* ncell_host, const double* box_host, int periodic, double rc, void* stream);
/* out[8] = n, ncells, stride, capacity per atom, max neighbours, ncell x, y, z */
int accu_nlist_info(accu_nlist* h, long long* out_host);
/* device int32 [n] copies of the cell id of each atom, the cell-sorted atom order and
 * the per-slot neighbour counts; any pointer may be NULL */
int accu_nlist_cells(accu_nlist* h, int* cell, int* sorted, int* nn, void* stream);
/* number of i < j pairs (synchronises), then the pairs themselves as device int64 */
int accu_nlist_count_half(accu_nlist* h, long long* npairs_host, void* stream);
int accu_nlist_export_half(accu_nlist* h, long long* i_out, long long* j_out, void* stream);
/* max_i |minimum image(pos_i - ref_i)|^2 (forces.py:90-93); synchronises */
int accu_max_disp2(const double* pos, const double* ref, int n, const double* box_host, int periodic,
                   double* out_host, void* stream);

/* ---- Lennard-Jones: forces.py:119-142 (list, h != NULL) and 157-175 (all pairs, h == NULL).
 * Species tables are host [ns][ns]: eps, sig^2, (cutoff*sig)^2; shift = src6^2 - src6.
 * sp: device int32 [n] species index. precision: ACCU_FP64 or ACCU_FP32_PAIR. */
int accu_lj_forces(accu_nlist* h, const double* pos, const int* sp, int n, int ns, const double* eps_host,
                   const double* sig2_host, const double* rc2_host, double shift, const double* box_host,
                   int periodic, int precision, double* F, double* U, void* stream);

/* ---- DampedShiftedCoulomb (shifted = 1): forces.py:408-452; Ewald real space
 * (shifted = 0): forces.py:371-379 plus the self term of forces.py:387.
 * erc = erfc(alpha * rcut) from the caller; ex_keys: device sorted int64 lo*n+hi of
 * excluded index pairs (nex may be 0). F and U are multiplied by k. */
int accu_coulomb_forces(accu_nlist* h, const double* pos, const double* q, int n, int shifted, double alpha,
                        double rcut, double erc, double k, const long long* ex_keys, int nex,
                        const double* box_host, int periodic, int precision, double* F, double* U, void* stream);

/* ---- Ewald reciprocal space: forces.py:380-386. cols_host: int32 [ncol][4] =
 * (nx, ny, nz_lo, nz_hi), the k-vectors of one half space; accumulate != 0 adds to F.
 * F += k * recip force, U = k * U_rec. */
int accu_ewald_recip(const double* pos, const double* q, int n, const double* box_host, double alpha,
                     const int* cols_host, int ncol, double k, int accumulate, double* F, double* U, void* stream);

/* ---- Gravity: forces.py:38-48; with targets (device int64 [ntgt]) the rows of
 * tree.py:81-87 direct_subset (F is then [ntgt][3]). origin (host [3], may be NULL)
 * is subtracted before float32 rounding in ACCU_FP32_PAIR mode. Targets must lie in
 * [0, n): the kernels read nothing for one outside and write its row of F as NaN. */
int accu_gravity_direct(const double* pos, const double* mass, int n, const long long* targets, int ntgt, double G,
                        double eps2, const double* box_host, int periodic, int precision, const double* origin_host,
                        double* F, double* U, void* stream);
/* The same computation without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Its energy and split partials live in a fixed
 * 8 MiB device array of the library (one per GPU, held until the process ends); calls
 * that need more return ACCU_EINVAL. The array keeps the last call's contents, so calls
 * that overlap on one GPU (two streams, or two threads) race and must be serialised by
 * the caller. Results equal accu_gravity_direct's bit for bit. */
int accu_gravity_direct_noalloc(const double* pos, const double* mass, int n, const long long* targets, int ntgt,
                                double G, double eps2, const double* box_host, int periodic, int precision,
                                const double* origin_host, double* F, double* U, void* stream);

/* ---- Barnes-Hut: tree.py:19-78 (BarnesHut._build and compute). The octree is
 * built and walked on the device; the handle keeps the tree and its buffers. */
typedef struct accu_bh accu_bh;
int accu_bh_create(accu_bh** out);
int accu_bh_destroy(accu_bh* h);
/* tree.py:30-53: the reference's octree (root cube, octant order, leaf_size, depth
 * < 48) over device pos [n][3] and m [n]; quadrupole != 0 also keeps traceless
 * quadrupole moments (an extension). Synchronises once, to size the node arrays. */
int accu_bh_build(accu_bh* h, const double* pos, const double* m, int n, int leaf_size, int quadrupole,
                  void* stream);
/* tree.py:54-78 over the last built tree: F [n][3] (user order) = m * acc and U =
 * 1/2 sum m phi (U may be NULL); quadrupole terms when the tree has them. */
int accu_bh_walk(accu_bh* h, double G, double eps2, double theta, double* F, double* U, void* stream);
/* out_host[6] = n, nodes, leaves, max depth, deep (1 if more than 21 levels were
 * sorted), leaf_size; synchronises */
int accu_bh_info(accu_bh* h, long long* out_host);
/* copies of the tree in pre-order (device buffers, any may be NULL): cube [nodes][4]
 * (center, half), com [nodes][4] (center of mass, mass), quad [nodes][6] (xx, yy, zz,
 * xy, xz, yz), next / leaf / start / count int32 [nodes] (first node after the
 * subtree, leaf flag, particle run in tree order), perm int32 [n] (tree order) */
int accu_bh_export(accu_bh* h, double* cube, double* com, double* quad, int* next, int* leaf, int* start,
                   int* count, int* perm, void* stream);

/* ---- device-resident N-body runs: Simulation.run (simulation.py:95-112) with
 * VelocityVerlet / Yoshida4 (integrators.py:19-47) over direct Gravity (forces.py:31-48)
 * or BarnesHut (tree.py:19-78). Host arrays in user order; box_host may be NULL (open
 * space) or a periodic box. Direct gravity fuses the force sum, the closing kick, the
 * record sums and the next opening kick and drift into one launch per substep. */
typedef struct accu_nbody accu_nbody;
int accu_nbody_create(accu_nbody** out, int n, const double* pos_host, const double* vel_host,
                      const double* mass_host, const double* box_host, int periodic);
int accu_nbody_destroy(accu_nbody* h);
/* Gravity (forces.py:31-48): precision ACCU_FP64, or ACCU_FP32_PAIR (open space; pair
 * terms relative to the initial centre of mass, sums in float64) */
int accu_nbody_set_direct(accu_nbody* h, double G, double eps2, int precision);
/* BarnesHut (tree.py:19-78), open space; quadrupole != 0 adds quadrupole terms */
int accu_nbody_set_bh(accu_nbody* h, double G, double eps2, double theta, int leaf_size, int quadrupole);
/* forces at t = 0 and record 0; log_cap records are kept */
int accu_nbody_init(accu_nbody* h, int log_cap);
/* steps steps of dt, each the substeps of weights w_host[0..nsub) (1 for VelocityVerlet;
 * w1, w0, w1 for Yoshida4); a record every record_every steps and at the last;
 * use_graph replays one captured hipGraph (direct gravity only) */
int accu_nbody_run(accu_nbody* h, int steps, double dt, const double* w_host, int nsub, int record_every,
                   int use_graph);
int accu_nbody_sync(accu_nbody* h);
/* state (host [n][3], any may be NULL): positions, velocities, accelerations F / m;
 * log_host [records][8] = PE, KE, Px, Py, Pz, Lx, Ly, Lz; synchronises */
int accu_nbody_get(accu_nbody* h, double* pos_host, double* vel_host, double* acc_host, double* log_host,
                   int log_max, long long* nrec_out);
int accu_nbody_stream(accu_nbody* h, void** stream_out);

/* ---- brick-tiled device-resident MD engine: Simulation.run (simulation.py:95-112) with
 * VelocityVerlet / Yoshida4 (integrators.py:19-47) over LJ (forces.py:119-142) and/or DSF
 * (forces.py:408-438), the Verlet rebuild rule of forces.py:86-98 applied on the device
 * every step and the record() sums of simulation.py:76-93 reduced in float64 on the
 * device, for one or many periodic systems at once, each with its own box, interactions and dt -
 * e.g. the dt, dt/2, dt/4 re-runs of timestep_convergence (accuracy.py:191-221).
 * Systems are given in one concatenated arena (user order); each needs
 * accu_bmd_set_system and the interaction calls before accu_bmd_init. */
typedef struct accu_bmd accu_bmd;
int accu_bmd_create(accu_bmd** out, int nsys, const int* natoms_host, const double* pos_host, const double* vel_host,
                    const int* sp_host, const double* box_host, int precision);
int accu_bmd_destroy(accu_bmd* h);
/* species masses/charges [ns], time step, list radius (incl. skin), skin, brick dims [3] */
int accu_bmd_set_system(accu_bmd* h, int sys, int ns, const double* mass_host, const double* charge_host, double dt,
                        double rc_list, double skin, const int* brick_dims_host);
int accu_bmd_set_lj(accu_bmd* h, int sys, int ns, const double* eps_host, const double* sig2_host,
                    const double* rc2_host, double shift);
int accu_bmd_set_dsf(accu_bmd* h, int sys, double alpha, double rcut, double erc, double k);
int accu_bmd_init(accu_bmd* h, double slack, int log_cap);
/* coarse steps of nticks ticks; at tick t the first active[t] systems advance one
 * (sub)step with closing-kick weight wc[t] and next-drift weight wn[t] (times each
 * system's dt); active[nticks-1] must be nsys; a record every record_every coarse steps */
int accu_bmd_run(accu_bmd* h, int coarse, int nticks, const int* active_host, const double* wc_host,
                 const double* wn_host, int record_every, int use_graph);
int accu_bmd_sync(accu_bmd* h);
/* state in user order; log_host [nsys][log_max][12] = PE, KE, Px, Py, Pz, rebuilds so
 * far, Lx, Ly, Lz (L = sum x cross m v over wrapped positions, system.py:112-114), 0, 0, 0
 * per record (Simulation.record, simulation.py:76-93); stats [nsys][4] (rebuilds, records);
 * info [16] = max neighbours, capacity, region capacity, region overflow, list entries,
 * then the last rebuild's phase ends in device wall-clock ticks after its start
 * ([5] binning, [6] scan, [7] grouping by cell (only when a cell outgrew its arrival
 * bucket), [8] = [7], [9] in-cell order and gather), [10] rebuilds in which a cell outgrew
 * its bucket, [11] the bucket capacity per cell, [12] the wall-clock rate in kHz */
int accu_bmd_get(accu_bmd* h, double* pos_host, double* vel_host, double* F_host, double* log_host, int log_max,
                 long long* stats_host, long long* info_host);
/* workgroup size of the per-brick kernels (multiple of 64 in [64, 512]; default 256) */
int accu_bmd_set_threads(accu_bmd* h, int threads);
/* device-resident trajectory frames (the positions/velocities that Simulation.record
 * appends, simulation.py:90-93, stored as io.py:19-25 saves them): a ring of cap_frames
 * frames of [N][3] positions then [N][3] velocities in user order, written by the fused
 * step of every `every`-th record (record 0 is the set-up state when this is called before
 * accu_bmd_init). cap_frames = 0 frees the ring. Frames beyond the capacity are counted,
 * not stored. */
int accu_bmd_set_snapshots(accu_bmd* h, int cap_frames, int every);
/* frames [first, first + count) to pos_host / vel_host ([count][N][3], either may be
 * NULL) and their record indices to rec_host [count] (may be NULL); taken_host: frames
 * taken so far (including those beyond the capacity) */
int accu_bmd_get_snapshots(accu_bmd* h, int first, int count, double* pos_host, double* vel_host, int* rec_host,
                           int* taken_host);
/* the snapshot ring in device memory, for the analysis entry points: *frames_out = frame 0
 * (each frame N positions then N velocities [N][3] in user order, 6 N doubles), *stored_out
 * = frames stored; valid until the ring is resized or the engine destroyed */
int accu_bmd_snapshot_frames(accu_bmd* h, void** frames_out, int* stored_out);
int accu_bmd_set_timing(accu_bmd* h, int enable);
int accu_bmd_timing(accu_bmd* h, double* total_ms_host, long long* launches_host);
/* integrator of every system of the arena: mode 0 velocity Verlet / Yoshida4 (the
 * accu_bmd_run weights), 1 Langevin BAOAB (integrators.py:50-71: temperature, friction
 * gamma), 2 Brownian (integrators.py:237-257: temperature, each species' diffusion from
 * accu_bmd_set_diffusion), 3 Nose-Hoover (integrators.py:74-116: temperature, period),
 * 4 Rattle (integrators.py:119-234, bonds from accu_bmd_set_constraints);
 * modes 1-4 run one step of dt per coarse step (accu_bmd_run with one tick, weights 1).
 * Langevin and Brownian noise: Philox4x32-10 normals keyed by seed, one counter per drift
 * of a system, one stream per atom user index. Nose-Hoover records carry the bath energy
 * (Q xi^2 / 2 + N_f kT eta, integrators.py:114-116) in log column 9. */
int accu_bmd_set_integrator(accu_bmd* h, int mode, double temperature, double friction, double period,
                            unsigned long long seed);
/* per-species diffusion coefficients D [ns] of system sys (Species.diffusion, system.py:17), before accu_bmd_init */
int accu_bmd_set_diffusion(accu_bmd* h, int sys, int ns, const double* D_host);
/* Nose-Hoover state per system to out_host [nsys][6]: xi, eta, Q, the velocity scales of the
 * last two half steps, k2 */
int accu_bmd_get_thermostat(accu_bmd* h, double* out_host);
/* Rigid bonds for integrator mode 4, Rattle (integrators.py:119-234): SHAKE after each drift
 * and RATTLE after each closing kick, both in float64, one thread per connected cluster of
 * bonds (the reference's dense M-SHAKE system is block diagonal over these clusters).
 * Before accu_bmd_init. Host arrays: cluster c has bonds [coff[c], coff[c+1]) with local
 * atom indices la / lb in [0, na_c) and lengths len (m), atoms [aoff[c], aoff[c+1]) with
 * global user indices uatom (system offsets included, each atom in at most one cluster),
 * system csys[c] (clusters ordered by system); at most 512 bonds per cluster. tol and
 * max_iter: the Newton stop |L^2 - r^2| / L^2 <= tol and iteration limit. */
int accu_bmd_set_constraints(accu_bmd* h, int nclusters, const int* coff, const int* aoff, const int* la,
                             const int* lb, const double* len, const int* uatom, const int* csys, double tol,
                             int max_iter);
/* out_host[0]: the worst relative bond-length error after any SHAKE so far (Rattle.max_violation),
 * out_host[1]: the residual at a failure; err_host[0]: 0 ok, 1 SHAKE did not converge in
 * max_iter, 2 SHAKE stalled (float64 cannot reach tol), 3 singular RATTLE system; err_host[1]: the failing
 * cluster's index c as passed to accu_bmd_set_constraints */
int accu_bmd_get_constraints(accu_bmd* h, double* out_host, int* err_host);

/* ---- analysis (accusim/analysis.py). pos / frames are device float64 arrays.
 * rdf pair counts (analysis.py:7-24): every pair i < j of pos [n][3] (minimum image when
 * periodic, box_host [3]) with |d| <= rmax, binned as numpy.histogram(r, bins, (0, rmax))
 * bins it, edges_host [bins + 1] being numpy.linspace(0, rmax, bins + 1); counts to
 * counts_host [bins] (bins <= 2048). */
int accu_rdf_counts(const double* pos, int n, const double* box_host, int periodic, double rmax,
                    const double* edges_host, int bins, unsigned long long* counts_host, void* stream);
/* msd (analysis.py:27-43) over nframes frames of n atoms, frame f at frames + f * stride
 * (doubles; [n][3] each), minimum-image steps when periodic; out_host [nframes] */
int accu_msd(const double* frames, int nframes, long long stride, int n, const double* box_host, int periodic,
             double* out_host, void* stream);
/* vacf (analysis.py:77-87): out_host [max_lag] = mean over atoms and origins t < nframes - k
 * of v(t) . v(t + k), velocity frames laid out as in accu_msd */
int accu_vacf(const double* frames, int nframes, long long stride, int n, int max_lag, double* out_host,
              void* stream);

/* ---- learned DSF -> Ewald pair correction (no counterpart in the reference; it
 * corrects DampedShiftedCoulomb, forces.py:391-452, toward Ewald, forces.py:333-388).
 * A 5 -> 64 -> 64 -> 1 pair MLP; parameters are 4609 float32 values laid out as
 * W1[64][5], b1[64], W2[64][64], b2[64], w3[64], b3. Layer 2 and the backward GEMMs
 * run on bf16 MFMA with float32 accumulation. */
typedef struct accu_mlp accu_mlp;
int accu_mlp_create(accu_mlp** out, const float* params_host);
int accu_mlp_destroy(accu_mlp* h);
/* copy the parameters out (get_host) and/or in (set_host); either may be NULL */
int accu_mlp_params(accu_mlp* h, float* get_host, const float* set_host);
/* pair cutoff rc, coordination switch r_on..r_off, coordination scale, force scale F0 (SI) */
int accu_mlp_configure(accu_mlp* h, double rc, double r_on, double r_off, double n_scale, double F0);
/* per-snapshot inputs: charges (device float64 times qscale) and coordination numbers
 * from the list h_nl (built with radius >= max(rc, r_off)) */
int accu_mlp_prepare(accu_mlp* h, accu_nlist* h_nl, const double* pos, const double* q, double qscale, int n,
                     const double* box_host, int periodic, void* stream);
/* correction forces (newtons) written to F, or added to it when accumulate != 0 */
int accu_mlp_forces(accu_mlp* h, accu_nlist* h_nl, const double* pos, int n, const double* box_host, int periodic,
                    int accumulate, double* F, void* stream);
/* gradient of the loss mean(((F_corr - target) / F0)^2) for one snapshot; target device
 * float64 [n][3]; grad_host [4609] and loss_host may be NULL (else this synchronises) */
int accu_mlp_grad(accu_mlp* h, accu_nlist* h_nl, const double* pos, int n, const double* box_host, int periodic,
                  const double* target, float* grad_host, double* loss_host, void* stream);
/* one Adam step with the last gradient */
int accu_mlp_adam(accu_mlp* h, double lr, double b1, double b2, double eps, void* stream);

#ifdef __cplusplus
}
#endif
#endif
