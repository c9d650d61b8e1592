/*
 * gravitas_hip.h -- C ABI of the MI355X (gfx950) Barnes-Hut backend.
 *
 * This library replaces the GPU bridge of the gravitas reference, i.e. the C
 * functions in the cgo preamble of internal/sim/bh/gpu_darwin.go:18-62 and
 * the Go stub internal/sim/bh/gpu_other.go:8-18, and moves the whole force
 * step bh.Solver.Accelerations (internal/sim/bh/bh.go:76-102) onto the GPU:
 * collect, Morton keys, stable radix sort, sorted copies, octree topology,
 * moments and the walk.  Plain pointers and sizes only; no C++ or torch types.
 *
 * Threading: one gx_solver is used by one thread at a time, as one bh.Solver
 * is (internal/sim/runner.go:177-226).  Different solvers may be used from
 * different threads.
 *
 * Error handling: functions returning int return GX_OK (0) or a GX_ERR_*
 * code; gx_error() returns the message of the calling thread's last error.
 */
#ifndef GRAVITAS_HIP_H
#define GRAVITAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1-3 keep the meaning of gpu_init's return values
 * (gpu_darwin.go:18-28: 1 no device, 2 library, 3 pipeline). */
#define GX_OK 0
#define GX_ERR_NO_DEVICE 1
#define GX_ERR_CODE_OBJECT 2
#define GX_ERR_PIPELINE 3
#define GX_ERR_ARG 4
#define GX_ERR_HIP 5
#define GX_ERR_CAPACITY 6
#define GX_ERR_UNSUPPORTED 7

/* Walk variants. */
#define GX_WALK_FP32 0 /* FP32 group walk, reference interaction set (default) */
#define GX_WALK_FP64 1 /* FP64 per-particle walk, bit-identical to the reference */
#define GX_WALK_FP32_Q16 2 /* FP32 group walk with FP16-stored quadrupoles (scaled by 1/(m h^2)) */

/* Body kinds, internal/sim/state.go:11-20. */
#define GX_KIND_STAR 0
#define GX_KIND_HOLE 1
#define GX_KIND_DARK 2
#define GX_KIND_GAS 3

/* Replaces gpu_init (gpu_darwin.go:18-28): selects and initialises the GPU,
 * returns GX_OK or GX_ERR_NO_DEVICE / GX_ERR_CODE_OBJECT and writes a message
 * to err.  Idempotent; the first result is kept, like GPUAvailable's
 * sync.Once (gpu_darwin.go:132-142). */
int gx_init(char *err, int errlen);

/* Replaces gpu_name (gpu_darwin.go:31-34): device name, "" before a
 * successful gx_init. */
void gx_name(char *out, int len);

/* Message of the calling thread's last error ("" if none). */
const char *gx_error(void);

/* A solver owns a device arena and a stream; buffers grow by 1.25x like the
 * Metal buffers (gpu_darwin.go:164-177) and are reused between calls. */
typedef struct gx_solver gx_solver;
gx_solver *gx_solver_new(void);
void gx_solver_free(gx_solver *s);

/* bh.Solver fields (bh.go:29-38): theta (0 means DefaultTheta 0.9, bh.go:69-74),
 * MonopoleOnly, and the walk variant (GX_WALK_*).  timing != 0 records
 * per-phase GPU times (gx_phase_ms). */
int gx_solver_config(gx_solver *s, double theta, int monopole_only, int walk, int timing);

/* Walk tuning (no counterpart in the reference): stars per walk group (32 or
 * 64; one wavefront per group) and the FP32 walk's work queue over resident
 * workgroups: 1 (the default) takes it when there are more groups than
 * resident workgroups, 0 never (one workgroup per group), 2 always. */
int gx_solver_tune(gx_solver *s, int group_size, int persistent);

/* Starting capacities (no counterpart in the reference): the node arena in
 * nodes and the FP32 walk's per-group lists (full-mask cells, partial-mask
 * cells, opened leaves), each > 0 to set, 0 to keep.  They grow on demand:
 * a step that overflows one is repeated with it enlarged, so small values
 * cost repeats, not results.  Buffers change size at the next step; a step
 * captured by gx_graph_capture_step must then be captured again. */
int gx_solver_reserve(gx_solver *s, int64_t nodes, int32_t full, int32_t partial, int32_t leaves);

/* Walk tuning with the FP32 walk's compile-time shapes (no counterpart in the
 * reference): as gx_solver_tune, plus the traversal stack (entries, 384 or
 * 768), the stars staged per evaluation flush (128, 256 or 384) and the
 * wavefronts per SIMD the registers are fitted for.  Built: 4 wavefronts with
 * the stack of 384 and staging 128, 256 or 384, and with the stack of 768 and
 * staging 256 or 384; 5 wavefronts with the stack of 384 and staging 128, 256
 * or 384; 6 with the stack of 384 and staging 256.  Other triples are
 * GX_ERR_ARG.  Every choice gives the same interaction set. */
int gx_solver_tune_walk(gx_solver *s, int group_size, int persistent, int stack, int parts, int waves);

/* One timed candidate of the walk autotuner. */
typedef struct {
    int32_t group_size, persistent, stack, parts, waves;
    float ms;                 /* mean GPU time of one walk (INFINITY: stack overflow) */
    int64_t pp_interactions;  /* the candidate's counts (equal for every candidate) */
    int64_t pc_interactions;
} gx_tune_entry;

/* Walk autotuner (no counterpart in the reference): on the given bodies
 * (device arrays and instances, as gx_batch_accelerations_dev) builds the
 * tree, then times reps FP32 walks of it for every candidate -- group size
 * 64 and 32, each walk shape of gx_solver_tune_walk, work queue
 * never (0) and always (2) -- and keeps the fastest in the solver.  The
 * choice is cached per workload shape (body count, instance count, walk,
 * monopole): a second call for the same shape applies it without timing.
 * Fills up to table_len entries of table and *entries with the candidate
 * count; ends with one force step of the chosen walk into ax, ay, az.  The
 * FP64 walk has nothing to tune (*entries = 0). */
int gx_solver_autotune(gx_solver *s, const double *x, const double *y, const double *z, const double *mass,
                       const uint8_t *kind, const uint8_t *alive, const int64_t *offsets, int32_t count,
                       const double *theta, const double *eps, int dark_matter_off, double *ax, double *ay,
                       double *az, int32_t reps, gx_tune_entry *table, int32_t table_len, int32_t *entries);

/* Uses the given hipStream_t (passed as void*) instead of the solver's own
 * stream; NULL restores the solver's stream. */
int gx_solver_set_stream(gx_solver *s, void *stream);

/* bh.Solver.Accelerations (bh.go:76-102) on host arrays of n bodies in state
 * order.  Star-star accelerations with Plummer softening eps; alive holes get
 * 0; dead bodies are left untouched; with dark_matter_off, dark matter pulls
 * nothing but is accelerated (sim.go:93-98).  Synchronous, like gpu_walk. */
int gx_accelerations(gx_solver *s, const double *x, const double *y, const double *z, const double *mass,
                     const uint8_t *kind, const uint8_t *alive, int64_t n, double eps, int dark_matter_off,
                     double *ax, double *ay, double *az);

/* Same on device arrays (hipMalloc'd or HBM-resident tensors).  Synchronous;
 * grows the node arena and repeats the step if the tree did not fit. */
int gx_accelerations_dev(gx_solver *s, const double *x, const double *y, const double *z, const double *mass,
                         const uint8_t *kind, const uint8_t *alive, int64_t n, double eps, int dark_matter_off,
                         double *ax, double *ay, double *az);

/* Enqueues the step on the solver's stream and returns at once; gx_sync
 * waits and reports GX_ERR_CAPACITY if the node arena was too small (the
 * outputs of such a step are not valid; gx_accelerations_dev regrows). */
int gx_enqueue_accelerations_dev(gx_solver *s, const double *x, const double *y, const double *z,
                                 const double *mass, const uint8_t *kind, const uint8_t *alive, int64_t n,
                                 double eps, int dark_matter_off, double *ax, double *ay, double *az);
int gx_sync(gx_solver *s);

/* KDK leapfrog on device arrays (internal/sim/sim.go:262-281 without holes
 * and gas; kick/drift :302-343 with c <= 0): if acc_valid is 0 the
 * accelerations are computed first (initForces, gas.go:14-26); then
 * v += dt/2 a, x += dt v, new accelerations, v += dt/2 a, over every body.
 * Synchronous. */
int gx_step_dev(gx_solver *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                const double *mass, const uint8_t *kind, const uint8_t *alive, int64_t n, double eps,
                int dark_matter_off, double dt, int acc_valid, double *ax, double *ay, double *az);

/* The two halves of the step on device arrays, enqueued on the solver's
 * stream: kick v += h a and drift x += dt v over n bodies. */
int gx_kick_dev(gx_solver *s, double *vx, double *vy, double *vz, const double *ax, const double *ay,
                const double *az, int64_t n, double h);
int gx_drift_dev(gx_solver *s, double *x, double *y, double *z, const double *vx, const double *vy,
                 const double *vz, int64_t n, double dt);

/* Batched force step (no counterpart in the reference, which runs one
 * Simulation per engine): count independent systems packed one after the
 * other in the device arrays; system b owns bodies [offsets[b], offsets[b+1])
 * (host array, offsets[0] = 0) and has its own theta[b] and eps[b] (host
 * arrays).  Each system gets exactly the accelerations a separate
 * gx_accelerations_dev call on its bodies would give; all systems are built
 * and walked in one pass (one walk launch).  Synchronous. */
int gx_batch_accelerations_dev(gx_solver *s, const double *x, const double *y, const double *z,
                               const double *mass, const uint8_t *kind, const uint8_t *alive,
                               const int64_t *offsets, int32_t count, const double *theta, const double *eps,
                               int dark_matter_off, double *ax, double *ay, double *az);

/* KDK leapfrog step of every system of a batch (as gx_step_dev). */
int gx_batch_step_dev(gx_solver *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                      const double *mass, const uint8_t *kind, const uint8_t *alive, const int64_t *offsets,
                      int32_t count, const double *theta, const double *eps, int dark_matter_off, double dt,
                      int acc_valid, double *ax, double *ay, double *az);

/* Captures one KDK step of the batch (kick, drift, force step, kick) into a
 * hipGraph on the solver's stream, after one eager force evaluation that
 * sizes every buffer; gx_graph_launch replays it `steps` times without host
 * work between steps, and gx_sync waits for the replays and reports an
 * overflow of any of them (GX_ERR_CAPACITY: the forces of that replay are
 * incomplete; run an eager step, which grows the buffers, and capture
 * again).  The arrays, the instance layout and the parameters are fixed at
 * capture; gx_graph_launch returns GX_ERR_ARG once a later eager call has
 * reallocated the buffers the graph uses. */
int gx_graph_capture_step(gx_solver *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                          const double *mass, const uint8_t *kind, const uint8_t *alive, const int64_t *offsets,
                          int32_t count, const double *theta, const double *eps, int dark_matter_off, double dt,
                          double *ax, double *ay, double *az);
int gx_graph_launch(gx_solver *s, int32_t steps);

/* Force step for an active set: as gx_accelerations_dev, but only bodies with
 * active[i] != 0 (device array, state order) get new accelerations; the tree
 * is still built from every alive star.  Each active star gets exactly the
 * value a full call would give it; the others keep theirs.  The reference's
 * counterpart is Hydro.Forces' active mask for gas (blockstep.go:166-172). */
int gx_accelerations_active_dev(gx_solver *s, const double *x, const double *y, const double *z,
                                const double *mass, const uint8_t *kind, const uint8_t *alive,
                                const uint8_t *active, int64_t n, double eps, int dark_matter_off, double *ax,
                                double *ay, double *az);

#define GX_BLOCK_MAX_LEVELS 8
#define GX_BLOCK_MAX_TICKS 128

typedef struct {
    int32_t levels;  /* rungs 0 .. levels-1 */
    int32_t events;  /* force evaluations in the step */
    int64_t bodies;  /* n */
    int64_t targets; /* stars evaluated, summed over the events */
    int64_t rung_count[GX_BLOCK_MAX_LEVELS]; /* bodies per rung at the start of the step */
    int32_t tick[GX_BLOCK_MAX_TICKS];        /* tick of each event (a tick is T / 2^(levels-1)) */
    int64_t active[GX_BLOCK_MAX_TICKS];      /* stars evaluated at each event */
} gx_block_stats;

/* One global step of length T with individual power-of-two time steps
 * (block steps), on device arrays.  The rung rules are those of the
 * reference's gas steps (internal/sim/blockstep.go:23-45, 81-90, event loop
 * :130-198): rung r steps by T/2^r, r <= levels-1, steps start only at
 * multiples of their length, every body drifts from event to event and the
 * bodies whose step ends are evaluated (gx_accelerations_active_dev), kicked
 * and given a new step.  A star's wanted step is sqrt(2 eta eps / |a|) (the
 * reference has no gravity step criterion); holes and dead bodies stay on
 * rung 0.  If acc_valid is 0 the forces are computed first.  The step ends
 * with every body synchronised and valid accelerations.  out (optional)
 * receives the rung populations and the active counts.  Synchronous. */
int gx_block_step_dev(gx_solver *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                      const double *mass, const uint8_t *kind, const uint8_t *alive, int64_t n, double eps,
                      int dark_matter_off, double T, int32_t levels, double eta, int acc_valid, double *ax,
                      double *ay, double *az, gx_block_stats *out);

/* One step with black holes, on device arrays: stepSplit
 * (internal/sim/split.go:164-199), the impulse/RESPA scheme.  Star-star forces
 * come from the tree once per step (the solver's walk); every force involving
 * an alive hole is a direct sum.  Stars whose timescale at a hole,
 * sqrt(r_eps / a), is below 20 steps dt, and all holes, take power-of-two
 * substeps (at most 256) from the shortest such timescale; far stars get the
 * holes' pull at the outer kicks.  Without alive holes it is step()'s KDK
 * (sim.go:266-277).  c is Params.C: <= 0 Newtonian; > 0 Paczynski-Wiita
 * radii for the pairs with holes and drifts with the proper velocity u/gamma
 * (relativity.go).  Restated for states without gas, with instant gravity
 * and non-spinning holes (gx_split_step_spin_dev adds the spins,
 * gx_split_step_retarded_dev retarded gravity).  If acc_valid is 0 the forces
 * are computed first.  *substeps (optional) receives the substep count
 * (Simulation.Substeps).
 * Synchronous. */
int gx_split_step_dev(gx_solver *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                      const double *mass, const uint8_t *kind, const uint8_t *alive, int64_t n, double eps,
                      int dark_matter_off, double dt, double c, int acc_valid, double *ax, double *ay, double *az,
                      int32_t *substeps);

/* stepSplit with spinning holes: as gx_split_step_dev, and spin_x/y/z (the
 * angular momentum J of every body, nonzero only for holes; sim.State.SpinX/Y/Z,
 * internal/sim/state.go:48-49) switch on frame dragging when c > 0
 * (internal/sim/framedrag.go): dragOuter after the first outer kick and
 * before the last one, dragFast around the drifts of every substep
 * (split.go:176-198).  The spins are updated in place. */
int gx_split_step_spin_dev(gx_solver *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                           const double *mass, const uint8_t *kind, const uint8_t *alive, double *spin_x,
                           double *spin_y, double *spin_z, int64_t n, double eps, int dark_matter_off, double dt,
                           double c, int acc_valid, double *ax, double *ay, double *az, int32_t *substeps);

/* Retarded gravity of black holes (internal/sim/retarded.go).  The modes are
 * sim.GravityMode (retarded.go:21-28): instant (Newtonian, infinite speed),
 * retarded (a body feels each hole where it was at the retarded time on the
 * body's past light cone) and retarded+v (that position extrapolated by the
 * hole's velocity times the delay).  They act only with c > 0, and only on
 * forces from holes; forces on holes stay instantaneous. */
#define GX_GRAVITY_INSTANT 0
#define GX_GRAVITY_RETARDED 1
#define GX_GRAVITY_RETARDED_V 2

/* The holes' histories of one simulation (Simulation.histories,
 * retarded.go:62-73): per alive hole, keyed by body ID, a ring buffer of its
 * position and coordinate velocity at the last 8,192 outer steps (384 KiB of
 * device memory each).  Used from one thread, with one solver at a time. */
typedef struct gx_retard gx_retard;

/* An empty history set (NULL if out of memory). */
gx_retard *gx_retard_new(void);
/* Frees the set and its device buffers; NULL is ignored. */
void gx_retard_free(gx_retard *r);
/* Number of histories held (len(Simulation.histories)). */
int gx_retard_count(const gx_retard *r);

/* holeAppeared (retarded.go:157-161; Simulation.AddBody of a hole,
 * sim.go:236-243): body i, a hole added at time, starts its history now with
 * its current state (device arrays) as the first entry, unless it has one; its
 * gravity spreads from there at speed c.  ids (device, may be NULL: the state
 * index) are the bodies' IDs; dt is Params.Dt.  Synchronous. */
int gx_retard_hole_appeared_dev(gx_solver *s, gx_retard *r, const double *x, const double *y, const double *z,
                                const double *vx, const double *vy, const double *vz, const uint64_t *ids, int64_t n,
                                int64_t i, double time, double dt, double c);

/* sourcePos (retarded.go:118-155) of alive hole number k (state order) for
 * the observers obs_x/y/z[0 .. m), all device arrays: where each sees the hole
 * at simulation time `time`, in src_x/y/z, and visible[j] = 0 where its gravity
 * has not arrived yet (src then 0).  With the histories of r bound as in a
 * step.  Synchronous. */
int gx_retard_source_pos_dev(gx_solver *s, gx_retard *r, const double *x, const double *y, const double *z,
                             const uint8_t *kind, const uint8_t *alive, const uint64_t *ids, int64_t n, int64_t k,
                             const double *obs_x, const double *obs_y, const double *obs_z, int64_t m, double time,
                             double dt, double c, int gravity, double *src_x, double *src_y, double *src_z,
                             uint8_t *visible);

/* stepSplit with retarded hole gravity: as gx_split_step_spin_dev (spin_x/y/z
 * may be NULL: no spins), and the pull of every hole on a body or on another
 * hole (holePull, split.go:120-137) comes from where the body sees it
 * (sourcePos, retarded.go:118-155) at simulation time `time`, from the
 * histories of r (r NULL: instant gravity).  After the step the alive holes
 * are recorded at time + dt (recordHoles, retarded.go:181-209, as Step does,
 * sim.go:285-287); outside the retarded modes (gravity instant or c <= 0) the
 * histories are dropped.  ids: device array of body IDs, NULL: the state
 * index.  Synchronous. */
int gx_split_step_retarded_dev(gx_solver *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                               const double *mass, const uint8_t *kind, const uint8_t *alive, double *spin_x,
                               double *spin_y, double *spin_z, const uint64_t *ids, int64_t n, double eps,
                               int dark_matter_off, double dt, double c, int acc_valid, double *ax, double *ay,
                               double *az, gx_retard *r, double time, int gravity, int32_t *substeps);

/* Simulation.dragOuter (internal/sim/framedrag.go:115-127) on device arrays:
 * if c > 0 and an alive hole spins, every alive body that is not a hole and
 * not near (near may be NULL: none) is kicked in state order by the
 * gravitomagnetic field of the holes' spins over time dt (a Boris rotation of
 * its proper velocity), and the holes take the opposite momentum and angular
 * momentum (velocities and spins updated in place).  Bit-identical to the
 * reference's sequential loop (see the notes in gx_drag.hip); with two or more
 * spinning holes the weights use r^2 sqrt(r^2) for math.Pow(r^2, 1.5).
 * eps is the softening.  Synchronous. */
int gx_drag_outer_dev(gx_solver *s, const double *x, const double *y, const double *z, double *vx, double *vy,
                      double *vz, const double *mass, const uint8_t *kind, const uint8_t *alive, const uint8_t *near,
                      double *spin_x, double *spin_y, double *spin_z, int64_t n, double eps, double dt, double c);

/* Simulation.dragFast (framedrag.go:129-143): the bodies near_list[0 ..
 * near_count) in that order (state indices; the near list of a split step),
 * then every alive hole in state order.  Synchronous. */
int gx_drag_fast_dev(gx_solver *s, const double *x, const double *y, const double *z, double *vx, double *vy,
                     double *vz, const double *mass, const uint8_t *kind, const uint8_t *alive,
                     const int32_t *near_list, int64_t near_count, double *spin_x, double *spin_y, double *spin_z,
                     int64_t n, double eps, double dt, double c);

typedef struct {
    int64_t passes;            /* passes over a list of bodies (dragOuter, dragFast's near part) */
    int64_t iterations;        /* fixed-point iterations over all passes */
    int64_t max_iterations;    /* most iterations of one pass */
    int64_t sequential_passes; /* passes done by one sequential thread instead */
    int64_t segments;          /* spin-chain segments scanned (1024 bodies each), over all iterations */
    int64_t slow_segments;     /* of these, added body by body */
} gx_drag_stats;

/* Frame-dragging counters since the solver was created (no counterpart in
 * the reference). */
int gx_drag_get_stats(const gx_solver *s, gx_drag_stats *out);

#define GX_DIAG_VALUES 11

/* FP64 diagnostics of internal/sim/diag.go on device arrays, summed on the
 * GPU in a fixed order: out[0] kinetic energy of the sources (sourceKinetic,
 * :68-77, c <= 0), out[1..3] momentum (Momentum :102-109), out[4..6] angular
 * momentum about the origin (AngularMomentum :112-119), out[7] total mass and
 * out[8..10] the mass-weighted position sum (CenterOfMass :122-134 before its
 * division).  Synchronous. */
int gx_diagnostics_dev(gx_solver *s, const double *x, const double *y, const double *z, const double *vx,
                       const double *vy, const double *vz, const double *mass, const uint8_t *kind, int64_t n,
                       int dark_matter_off, double *out);

/* bh.Solver.PotentialEnergy (potential.go:12-42, holePotential :84-96) on
 * device arrays: theta capped at 0.9; FP64 tree walk for star-star pairs and
 * exact pair sums (PairPotential, pair.go:37-47) for the pairs with alive
 * holes; c is Params.C (<= 0: Newtonian, else Paczynski-Wiita radii for
 * star-hole pairs).  Synchronous. */
int gx_potential_energy_dev(gx_solver *s, const double *x, const double *y, const double *z, const double *mass,
                            const uint8_t *kind, const uint8_t *alive, int64_t n, double eps, int dark_matter_off,
                            double c, double *out);

/* PotentialEnergyParallel (internal/sim/diag.go:44-64): the exact potential
 * energy of the pair model (pair.go) over every pair of the n bodies (alive
 * or not, as the reference's loop), FP64 on the GPU: softened Newtonian
 * pairs, Paczynski-Wiita star-hole pairs for c > 0, dark matter with
 * dark_matter_off pulling nothing.  TotalEnergy (diag.go:31-41) uses it up to
 * 30,000 bodies and the tree estimate (gx_potential_energy_dev) above.  The
 * rows' sums are added in state order (the reference adds chunk sums in
 * completion order). */
int gx_potential_exact_dev(gx_solver *s, const double *x, const double *y, const double *z, const double *mass,
                           const uint8_t *kind, int64_t n, double eps, int dark_matter_off, double c, double *out);

/* FP64 direct sum (internal/sim/direct/direct.go:25-48) for nt target
 * indices (device int64) against all n bodies (device arrays); out is a
 * device array of nt*3 doubles (ax, ay, az per target).  Synchronous. */
int gx_direct_sample_dev(gx_solver *s, const double *x, const double *y, const double *z, const double *mass,
                         const uint8_t *kind, int64_t n, const int64_t *targets, int64_t nt, double eps,
                         int dark_matter_off, double *out);

/* Force step without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid allocation).  As gx_accelerations_dev on device
 * arrays of at most GX_NOALLOC_MAX_BODIES bodies, with theta (0: 0.9), the
 * softening and the walk (GX_WALK_*) given per call, launched on `stream`
 * (a hipStream_t as void*, NULL: the default stream) and synchronous.  Every
 * buffer of the step is a slice of one static device array of
 * GX_NOALLOC_SCRATCH_BYTES per GPU, which each GPU the library runs on holds
 * until the process ends; no hipMalloc, hipFree or other allocating call is
 * made.  The FP32 walk's lists have fixed capacities: a group that overflows
 * them is walked again by the FP64 walk (same interaction set).  A tree that
 * needs more nodes than the scratch holds is GX_ERR_CAPACITY; more than
 * GX_NOALLOC_MAX_BODIES bodies is GX_ERR_ARG.  Calls share their GPU's
 * scratch, and each leaves its contents for the next: calls on one GPU are
 * serialised (a later call waits for an earlier one); work that overlaps a
 * call on the same GPU and touches its scratch races with it. */
#define GX_NOALLOC_MAX_BODIES 65536
#define GX_NOALLOC_SCRATCH_BYTES (34u << 20)
int gx_accelerations_noalloc(const double *x, const double *y, const double *z, const double *mass,
                             const uint8_t *kind, const uint8_t *alive, int64_t n, double theta, double eps,
                             int dark_matter_off, int walk, double *ax, double *ay, double *az, void *stream);

/* LastGPUSeconds (gpu_darwin.go:155): GPU time of the last force step
 * (build and walk, from its first kernel to its last; after gx_graph_launch,
 * the mean per replay), recorded on every step whatever the timing setting;
 * 0 before the first step. */
double gx_last_seconds(const gx_solver *s);

typedef struct {
    int64_t n_stars, n_holes, n_nodes;
    int64_t pp_interactions; /* star-star pairs, self pairs excluded */
    int64_t pc_interactions; /* star-cell pairs */
    int32_t max_stack, flags, depth, walk;
    int64_t fallbacks;       /* FP32 steps recomputed by the FP64 walk */
    int64_t list_entries;    /* FP32 walk: cell-list entries over all groups */
    int64_t rounds;          /* FP32 walk: traversal rounds over all groups */
    int64_t mixed_nodes;     /* FP32 walk: nodes decided star by star */
    int32_t max_list[4];     /* FP32 walk: longest full-cell, partial-cell, leaf list; unused */
    int64_t n_targets;       /* stars whose forces were computed (the active ones of an active set) */
} gx_stats;

/* Statistics of the last completed step. */
int gx_get_stats(const gx_solver *s, gx_stats *out);

/* Per-phase GPU milliseconds of the last step when timing is on:
 * 0 collect (which also writes the Morton keys), 1 keys (empty since), 2 sort,
 * 3 sorted copies, 4 topology, 5 moments, 6 walk (for the FP32 walk run as
 * two kernels: its traversal kernel), 7 total, 8 FP32 walk evaluation kernel
 * (0 unless the walk runs as two kernels).  Returns the number of phases
 * written. */
int gx_phase_ms(const gx_solver *s, float *ms, int n);

/* Test hook with no counterpart in the reference: copies count items of
 * the last build to host memory.  what: 0 keys (uint64, sorted), 1 order
 * (int32, sorted position -> state index), 2 sorted stars (4 doubles x, y, z,
 * source mass), 3 nodes (64-byte records: double x, y, z, crit2; int32 first,
 * next, lo, hi, nchild, level, parent, pad), 4 FP64 moments (8 doubles m, qxx,
 * qyy, qzz, qxy, qxz, qyz, pad).  Returns the number of items copied. */
int64_t gx_export(gx_solver *s, int what, void *host, int64_t count);

/* ---------------------------------------------------------------------- */
/* SPH hydro: neighbour BVH and forces (sim.Hydro, internal/sph)          */

/* sph.Solver (internal/sph/sph.go:11-19): owns the neighbour tree and the
 * scratch of the hydro on the GPU, and its own stream.  NULL without a
 * device (gx_error tells why).  One gx_hydro is used by one thread at a
 * time, as one sph.Solver is; a hydro and a solver may run side by side. */
typedef struct gx_hydro gx_hydro;
gx_hydro *gx_hydro_new(void);
void gx_hydro_free(gx_hydro *h);
/* timing != 0 records per-phase GPU times (gx_hydro_phase_ms). */
int gx_hydro_config(gx_hydro *h, int timing);
int gx_hydro_set_stream(gx_hydro *h, void *stream);

/* The fields of sim.GasParams the hydro reads (internal/sim/sim.go:26-29);
 * 0 takes the defaults 5/3, 40, 1 (sph.go:64-76). */
typedef struct {
    double gamma, neighbors, alpha;
} gx_gas_params;

/* sph.Solver.Forces (internal/sph/sph.go:83-263) on device arrays of n
 * bodies in state order.  Alive bodies of kind GX_KIND_GAS take part.  h
 * (smoothing lengths; <= 0: first guess, sph.go:277-302) and rho (densities)
 * are read and updated in place; u is the internal energy.  active (may be
 * NULL: all) selects the evaluated bodies, step (may be NULL) their current
 * time steps for the limiter.  For evaluated bodies ax, ay, az (pressure and
 * viscosity acceleration), dudt, divv and dt are written; wake holds float64
 * bits and is lowered atomically for neighbours whose step is too long
 * (sim.HydroOut, internal/sim/sim.go:60-68).  FP64, bit-identical to the
 * reference's order of operations.  Synchronous. */
int gx_hydro_forces_dev(gx_hydro *h, const double *x, const double *y, const double *z, const double *vx,
                        const double *vy, const double *vz, const double *mass, const uint8_t *kind,
                        const uint8_t *alive, const double *u, double *hsml, double *rho, int64_t n,
                        const gx_gas_params *gas, double softening, const uint8_t *active, const double *step,
                        double *ax, double *ay, double *az, double *dudt, double *divv, double *dt,
                        uint64_t *wake);

/* Same on host arrays (copies in and out), the drop-in for the Go engine's
 * Hydro.Forces call (internal/sim/gas.go:14-26). */
int gx_hydro_forces(gx_hydro *h, const double *x, const double *y, const double *z, const double *vx,
                    const double *vy, const double *vz, const double *mass, const uint8_t *kind,
                    const uint8_t *alive, const double *u, double *hsml, double *rho, int64_t n,
                    const gx_gas_params *gas, double softening, const uint8_t *active, const double *step,
                    double *ax, double *ay, double *az, double *dudt, double *divv, double *dt, uint64_t *wake);

typedef struct {
    int64_t n_gas, n_nodes, n_leaves;
    int32_t depth, max_iterations;  /* tree levels; most density passes of one body */
    int64_t density_sums;           /* density evaluations over all bodies and passes */
    int64_t pairs;                  /* force-pass neighbour pairs (j != i) */
    int64_t walk_steps;             /* walk records the force pass's wavefronts loaded */
    int64_t density_pairs;          /* neighbour terms summed by the density passes (self included) */
    int64_t visc_pairs;             /* force pairs that approach (the viscosity branch) */
} gx_hydro_stats;

int gx_hydro_get_stats(const gx_hydro *h, gx_hydro_stats *out);

/* Per-phase GPU milliseconds of the last evaluation when timing is on:
 * 0 gas list, keys and sort, 1 tree build and refit, 2 density passes,
 * 3 leaf sort and per-body terms, 4 force pass, 5 total.  Returns the
 * number written. */
int gx_hydro_phase_ms(const gx_hydro *h, float *ms, int n);

/* Test hook with no counterpart in the reference: copies count items of the
 * last evaluation's tree to host memory.  what: 0 keys (uint64, sorted), 1
 * order (int32, tree position -> state index, leaves sorted by decreasing
 * h), 2 nodes (80-byte records: double minx, miny, minz, maxx, maxy, maxz,
 * hmax; int32 lo, hi, first (right child), second (left child), escape (the
 * next node after the subtree in a right-first walk, -1 at the end), parent
 * (-1 at the root)).  Returns the number of items copied. */
int64_t gx_hydro_export(gx_hydro *h, int what, void *host, int64_t count);

#ifdef __cplusplus
}
#endif
#endif
