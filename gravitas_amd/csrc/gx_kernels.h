// gx_kernels.h -- kernel declarations shared between the translation units.
#pragma once
#include "gx_common.h"

namespace gx {

// gx_api.hip: the thread's error message (gx_error) and the device check of
// gx_init, for entry points defined in other files.
int set_error(int code, const char* fmt, ...);
int device_ready();

constexpr int kCollectTile = 1024;  // bodies per collect workgroup (256 threads x 4)
constexpr int kSortTile = 2048;     // keys per sort tile (256 threads x 8; the scatter stages a tile in LDS)
constexpr int kTopoTile = 1024;     // stars per topology tile (256 threads x 4)

// gx_build.hip
__global__ void k_step_init(Meta* meta, Batch batch);
__global__ void k_collect_count(const double* x, const double* y, const double* z, const uint8_t* kind,
                                const uint8_t* alive, long long n, int* blockStar, int* blockHole, double* blockBox,
                                Batch batch);
__global__ void k_collect_scan(int nb, int* blockStar, int* blockHole, const double* blockBox, Meta* meta,
                               Batch batch);
__global__ void k_collect_scatter(const double* x, const double* y, const double* z, const uint8_t* kind,
                                  const uint8_t* alive, long long n, const int* starOff, const int* holeOff,
                                  Batch batch, int32_t* order, int32_t* holes, uint64_t* keys, uint16_t* instOf);
__global__ void k_sort_hist(const uint64_t* keys, const int32_t* vals, const Meta* meta, int shift, int instPass,
                            Batch batch, uint32_t* hist, int ntiles);
constexpr int kScanThreads = 1024;  // threads of the one-workgroup scans (collect, active sets)
constexpr int kRowScanThreads = 256;  // threads of the per-digit scan over the tiles
__global__ void k_sort_scan(uint32_t* hist, int ntiles, uint32_t* digitTotal);
__global__ void k_sort_scatter(const uint64_t* kin, const int32_t* vin, uint64_t* kout, int32_t* vout,
                               const Meta* meta, int shift, int instPass, Batch batch, const uint32_t* hist,
                               const uint32_t* digitTotal, int ntiles, int32_t* rank);
__global__ void k_scatter_spos(const double* x, const double* y, const double* z, const double* m,
                               const uint8_t* kind, const uint8_t* alive, long long n, const int32_t* rank,
                               int dmOff, double4* spos);
__global__ void k_topo_count(const uint64_t* keys, Meta* meta, Batch batch, uint16_t* chain,
                             uint32_t* tileCount, int ntiles);
__global__ void k_topo_nodes(const uint64_t* keys, const uint16_t* chain, const uint32_t* tileOff,
                             const uint32_t* levelTotal, Meta* meta, Batch batch, TravNode* trav, int32_t* chainFirst,
                             long long cap, int ntiles);
__global__ void k_topo_links(const uint64_t* keys, const uint16_t* chain, TravNode* trav, const int32_t* chainFirst,
                             Meta* meta, Batch batch);
__global__ void k_moments_level(int L, const double4* spos, TravNode* trav, Mult64* m64, Mult32* m32,
                                Mult16* m16, const Meta* meta, Batch batch);
__global__ void k_moments_leaves(const double4* spos, TravNode* trav, Mult64* m64, Mult32* m32, Mult16* m16,
                                 Meta* meta, Batch batch, int Lmax);
__global__ void k_zero_holes(const int32_t* holes, Meta* meta, double* ax, double* ay, double* az, int accumulate);
__global__ void k_kick(long long n, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                       const double* az, double h);
__global__ void k_drift(long long n, double* px, double* py, double* pz, const double* vx, const double* vy,
                        const double* vz, double dt);
__global__ void k_active_count(const uint8_t* active, const int32_t* order, const Meta* meta, uint8_t* actS,
                               int* blockAct);
__global__ void k_active_scan(int nb, int* blockAct, Meta* meta);
__global__ void k_active_inst(const uint8_t* actS, const int* blockAct, const Meta* meta, Batch batch, long long* tOff);
__global__ void k_active_groups(Batch batch, const long long* tOff);
__global__ void k_active_scatter(const uint8_t* actS, const int* blockAct, const Meta* meta, int32_t* tlist);
__global__ void k_block_active(long long n, const int8_t* rung, int now, int full, uint8_t* active);
__global__ void k_block_kick(long long n, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                             const double* az, const uint8_t* kind, const uint8_t* alive, int8_t* rung, int now,
                             int full, int maxRung, double T, double crit, unsigned long long* hist);
__global__ void k_fill_u64(unsigned long long* p, int count, unsigned long long v);
__global__ void k_walk_reset(Meta* meta);
constexpr int kDiagValues = 11;
constexpr int kDiagBlocks = 1024;
__global__ void k_diag(long long n, const double* px, const double* py, const double* pz, const double* vx,
                       const double* vy, const double* vz, const double* m, const uint8_t* kind, int dmOff,
                       double* partial);

// gx_walk64.hip
__global__ void k_walk64(const double4* spos, const TravNode* trav, const Mult64* m64, const int32_t* order,
                         Meta* meta, Batch batch, int quad, double* ax, double* ay, double* az);
__global__ void k_potential64(const double4* spos, const TravNode* trav, const Mult64* m64, Meta* meta, Batch batch,
                              int quad, const int32_t* order, const int32_t* holes, const double* px, const double* py,
                              const double* pz, const double* mass, const uint8_t* kind, int dmOff, double c,
                              double* partial);
__global__ void k_hole_potential(const double4* spos, const Meta* meta, const int32_t* holes, const double* px,
                                 const double* py, const double* pz, const double* mass, double eps2, double c,
                                 double* out);
__global__ void k_direct_sample(const double* x, const double* y, const double* z, const double* m,
                                const uint8_t* kind, long long n, const long long* targets, int nt, double eps2,
                                int dmOff, int splits, double* partial);
__global__ void k_direct_reduce(const double* partial, int nt, int splits, double* out);
__global__ void k_potential_exact(const double* x, const double* y, const double* z, const double* m,
                                  const uint8_t* kind, long long n, double eps2, int dmOff, double c, int slices,
                                  double* partial);

// gx_drag.hip: frame dragging (internal/sim/framedrag.go)
struct DragArgs {
    const double *x, *y, *z, *mass;
    const uint8_t* alive;
    const int32_t* holes;  // the alive holes in state order
    int nh;
    double eps2, c, dt;
};
// Integer summary of one segment of an exact ordered scan.
struct DragSeg {
    long long total, pmin, pmax;  // sum of the integer steps; bounds of the partial sums (0 included)
    long long qf;                 // sum of the integer steps up to the first tie (kSegHasTie)
    int e, flags;                 // binade exponent guessed for the segment; kSeg* flags
};
constexpr int kDragSeg = 1024;  // elements per segment
constexpr int kDragRow = 64;    // elements per speculated row
constexpr int kDragRows = kDragSeg / kDragRow;
constexpr int kSegNoGuess = 1, kSegBig = 2, kSegHasTie = 4, kSegAllNegZero = 8;
__global__ void k_drag_spinning(const int32_t* holes, const Meta* meta, const double* jx, const double* jy,
                                const double* jz, double c, int* on);
__global__ void k_drag_outer_flag(long long n, const uint8_t* kind, const uint8_t* alive, const uint8_t* nearFlag,
                                  uint8_t* flag);
__global__ void k_drag_chain_io(const int32_t* holes, int nh, double* ax, double* ay, double* az, double* vals,
                                int writeBack);
__global__ void k_drag_eval(DragArgs a, const int32_t* list, long long ne, double* vx, double* vy, double* vz,
                            const double* jx, const double* jy, const double* jz, const double* S, int apply,
                            const int* on, double* D);
__global__ void k_drag_seq(DragArgs a, const int32_t* list, long long ne, double* vx, double* vy, double* vz,
                           double* jx, double* jy, double* jz, const int* on);
constexpr int kDragSmall = 4096;  // lists up to this length run in one workgroup (k_drag_small)
__global__ void k_drag_small(DragArgs a, const int32_t* list, long long ne, double* vx, double* vy, double* vz,
                             double* jx, double* jy, double* jz, const int* on, double* D, double* S0, double* S1,
                             double* E, int maxIt, int holesToo, unsigned long long* counters);
__global__ void k_scan_segsum(const double* D, long long ne, int chains, int nseg, double* sig, double* rowSig);
__global__ void k_scan_guess(const double* sig, int nseg, const double* start, double* guess);
__global__ void k_scan_summary(const double* D, long long ne, int chains, int nseg, const double* guess,
                               DragSeg* seg);
__global__ void k_scan_spec(const double* D, long long ne, int chains, int nseg, const double* start,
                            const double* guess, const double* rowSig, const double* Sold, double* S,
                            double* specStart, double* specEnd, int* specOk);
__global__ void k_scan_chain(const double* D, long long ne, int nseg, const DragSeg* seg, const double* start,
                             const double* specStart, const double* specEnd, const int* specOk, double* S,
                             double* segStart, int* segMode, double* end, unsigned long long* counters);
__global__ void k_scan_fill(const double* D, long long ne, int chains, int nseg, const DragSeg* seg,
                            const double* segStart, const int* segMode, const double* start, const double* Sold,
                            double* S, int* changed);

// gx_split.hip
// Retarded gravity of black holes (internal/sim/retarded.go).  holeHistory
// (:62-73) as the kernels read it: the header, kept by the host, and a ring of
// kHistLen entries per coordinate in device memory (x, y, z, vx, vy, vz back
// to back; the velocities are coordinate velocities u / gamma).
constexpr int kHistLen = 8192;      // historyLen, retarded.go:50-52
constexpr double kRetardTol = 1e-7;  // retardTolerance, :54-59
constexpr int kRetardMaxIt = 64;    // retardMaxIterations
struct HistDev {
    const double* ring;  // nullptr: the hole has no history (bindHistories found none)
    double t0, appeared, dt;
    int n, head;
};
// The holes' sources during one step (sourcePos, retarded.go:118-155).
struct RetardDev {
    const HistDev* hist;  // per hole slot (m.holes order); nullptr: instant gravity
    int mode;             // 1 GravityRetarded, 2 GravityRetardedExtrapolated
    double time, dt;      // Simulation.Time during the step, Params.Dt
};
// One push (holeHistory.push, retarded.go:75-83) of body `body` into slot
// `pos` of a ring.
struct HistPush {
    double* ring;
    int32_t body, pos;
};
__global__ void k_retard_push(const HistPush* p, int cnt, const double* x, const double* y, const double* z,
                              const double* vx, const double* vy, const double* vz, double c);
__global__ void k_retard_source(RetardDev r, long long k, const int32_t* holes, const double* x, const double* y,
                                const double* z, const double* ox, const double* oy, const double* oz, long long m,
                                double c, double* sx, double* sy, double* sz, uint8_t* visible);
__global__ void k_gather_hole_ids(const int32_t* holes, long long nh, const uint64_t* ids, uint64_t* out);
__global__ void k_split_classify(long long n, const double* x, const double* y, const double* z, const double* mass,
                                 const uint8_t* kind, const int32_t* holes, const Meta* meta, double eps2, double c,
                                 double limit, uint8_t* nearFlag, unsigned long long* tauEnc);
__global__ void k_flag_count(long long n, const uint8_t* flag, int* tileCnt);
__global__ void k_flag_scan(int nb, int* tileCnt, long long* total);
__global__ void k_flag_scatter(long long n, const uint8_t* flag, const int* tileOff, int32_t* list);
__global__ void k_split_far(long long n, const double* x, const double* y, const double* z, const double* mass,
                            const uint8_t* kind, const uint8_t* nearFlag, const int32_t* holes, const Meta* meta,
                            double eps2, double c, RetardDev r, double* fx, double* fy, double* fz);
__global__ void k_split_fast(const int32_t* nearList, const long long* nn, const double* x, const double* y,
                             const double* z, const double* mass, const uint8_t* kind, const int32_t* holes,
                             const Meta* meta, double eps2, double c, RetardDev r, double* fx, double* fy,
                             double* fz);
__global__ void k_split_hole_sum(long long n, const double* x, const double* y, const double* z, const double* mass,
                                 const uint8_t* kind, int dmOff, const uint8_t* nearFlag, const int32_t* nearList,
                                 const long long* nn, const int32_t* holes, const Meta* meta, double eps2, double c,
                                 int nearMode, double* partial);
__global__ void k_split_hole_finish(const double* x, const double* y, const double* z, const double* mass,
                                    const uint8_t* kind, const int32_t* holes, const Meta* meta, double eps2, double c,
                                    RetardDev r, int nearMode, int slices, const double* partial, double* fx,
                                    double* fy, double* fz);
__global__ void k_split_kick_outer(long long n, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                                   const double* az, const double* fx, const double* fy, const double* fz, double h);
__global__ void k_split_kick_list(const int32_t* list, long long cntHost, const long long* cntDev, double* vx,
                                  double* vy, double* vz, const double* fx, const double* fy, const double* fz,
                                  double h);
__global__ void k_split_drift_list(const int32_t* list, long long cntHost, const long long* cntDev, double* x,
                                   double* y, double* z, const double* vx, const double* vy, const double* vz,
                                   double dt, double c);
__global__ void k_split_drift_far(long long n, const uint8_t* kind, const uint8_t* nearFlag, double* x, double* y,
                                  double* z, const double* vx, const double* vy, const double* vz, double dt, double c);

// gx_walk32.hip
// Per-group interaction lists written by the FP32 walk's traversal and read by
// its evaluation: full-mask cells as bare node indices, partial-mask cells and
// opened leaves as (node, mask lo, mask hi).  A region holds one group's
// lists: leaves opened by every star fill its leaf list from the front, the
// others from the back.  Region r owns entries [r*cap, (r+1)*cap) of each
// list; with perSlot the region is the workgroup's (the work queue walks one
// group at a time per workgroup, so the lists of all groups need only one
// region per resident workgroup), otherwise the group's.
struct GroupCounts {
    int nf, np;      // full-mask and partial-mask cells
    int nlf, nlp;    // leaves opened by every star, by some
    int nps;         // stars in the leaves opened by some
    int overflow;    // the traversal stack overflowed
    int pad[2];
};
// Largest per-group list capacity (entries are addressed with 24-bit products).
constexpr int kMaxListCap = (1 << 24) - 1;

struct Lists {
    uint32_t* full;
    uint3* part;
    uint3* leaf;
    GroupCounts* counts;
    int cf, cp, cl;  // capacities per region
    int perSlot;
};

// Compile-time shapes of the FP32 walk: traversal stack entries, stars staged
// per evaluation flush, and the wavefronts per SIMD the register allocation
// must leave room for.  4 is what the walk needs unconstrained; 5 and 6
// make the compiler fit 96 and 80 VGPRs (at 6 without the cell prefetch and
// with a few spills), which only pays where the LDS of a workgroup lets that
// many stay resident: the stack of 384 with the staging of 128, 256 or 384
// (4.75, 5.25 and 7.4 KiB, up to 32, 30 and 21 wavefronts per CU).
// kWalkConfigs lists the instances; the autotuner picks among them
// (gx_solver_autotune).
struct WalkConfig {
    int stack, parts, waves;
};
constexpr int kWalkConfigs = 9;
constexpr WalkConfig kWalkConfig[kWalkConfigs] = {{384, 128, 4}, {384, 256, 4}, {384, 384, 4},
                                                  {384, 256, 6}, {768, 256, 4}, {768, 384, 4},
                                                  {384, 128, 5}, {384, 384, 5}, {384, 256, 5}};
constexpr int kWalkDefault = 7;  // 384 entries, 384 stars, 5 waves

// m16 non-null selects the FP16-stored quadrupoles; eps0 marks a batch with an
// unsoftened instance; persistentBlocks > 0 runs the fused walk as a work queue
// over that many resident workgroups; config indexes kWalkConfig.
void launch_walk32(const double4* spos, const TravNode* trav, const Mult32* m32, const Mult16* m16,
                   const int32_t* order, Meta* meta, const Batch& batch, long long ngroupsUpper, bool eps0,
                   int quad, const Lists& lists, double* ax, double* ay, double* az, hipStream_t stream,
                   hipEvent_t mid, bool split, int persistentBlocks, int config);
// Resident workgroups per CU of the fused FP32 walk of a configuration (for
// the persistent grid).
int walk32_blocks_per_cu(int config);

}  // namespace gx
