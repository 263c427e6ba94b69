// mgp_batch.h -- the batched fast solver: many independent problems of
// uneven sizes, each solved as KRYLOV_MG solves it, with one launch per level
// operation for all of them.  Launches are driven by tables built once per
// batch (row tasks for the line relaxation, cell tiles for everything else);
// every workgroup reads its problem's entry and works on that problem alone,
// so a problem's arithmetic does not depend on the others in the batch.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "mgp_common.h"
#include "mgp_host.h"

namespace mgp {

// One level of one problem as the batched level kernels see it.  e/f are the
// correction and right-hand side of the level's cycle.  On the finest level
// the two entries are the two preconditioner applications of an iteration
// (phat from p, shat from s); on the others both are the level's own fields.
struct BLev {
    Grid g;
    Geo geo;
    double *e[2];
    double *f[2];
    const double *K;    // ring stencil table
    double slo, shi;    // ghost signs of the y sides in the prolongation
    int prob, nt;       // problem index; x-line chunks (nx / C)
    int off;            // the fine row of the next coarser level's row 0 (semicoarsen_step)
    int sweeps;         // coarsest level of its problem: zebra sweeps there (else 0)
    // FP32 copies for the mixed-precision cycle (null otherwise): metrics,
    // stencil table, correction and right-hand side (one pair: the two
    // applications of an iteration run one after the other)
    GeoF geo32;
    const float *K32;
    float *e32, *f32;
};

// Krylov vectors of one problem (finest level, level layout).
enum BVec : int { BV_PHI = 0, BV_RHS, BV_R, BV_RHAT, BV_P, BV_V, BV_S, BV_T, BV_PHAT, BV_SHAT, BV_COUNT };

struct BProb {
    Grid g;
    Geo geo;
    Bc bc;             // the caller's BCs (homogeneous use)
    double *v[BV_COUNT];
    double *sc;        // SC_COUNT BiCGStab scalars
    double *partial;   // 2 nb partial sums of the reductions (two products)
    int nb;            // reduction row blocks: block b sums rows b, b + nb, ...
    int lev0;          // index of the finest level in the level table
};

// Row task: row `row` of level `lev`, then row `extra` (< 0: none) by the
// same workgroup (the boundary-row fold of the single-problem sweep).
struct XlTask {
    int lev, row, extra, pad;
};

// Cell tile: cells (64 tx .. 64 tx + 63, BT_ROWS ty .. BT_ROWS ty + 15) of
// item `item` (a problem or a level, as the kernel says), each thread of a
// 256-thread workgroup taking four rows 4 apart; for reductions, tx is the
// row block.
constexpr int BT_ROWS = 16;
struct BTile {
    int item, tx, ty, pad;
};

// Level kernels (mgp_xline.hip).  `active` is indexed by problem; sel picks
// the e/f pair of the finest level.
// k >= 0: sweep k of a level's coarsest-level sweeps (levels with fewer return)
// fp32: the FP32 fields of the levels (e32/f32, geo32, K32) instead of e[sel]/f[sel]
void launch_xline_batch(hipStream_t s, const XlTask *tasks, int ntasks, const BLev *levs, const int *active, int sel,
                        int C, int threads, size_t lds, int k, bool fp32);
// coarse (lev+1) r = 1x2 restriction of f - L e on lev, coarse e = 0; tiles over the coarse cells
void launch_restrict_batch(hipStream_t s, const BTile *tiles, int ntiles, const BLev *levs, const int *active,
                           int sel, bool fp32);
// e(lev) += linear-in-y prolongation of e(lev+1); tiles over the fine cells
void launch_prolong_batch(hipStream_t s, const BTile *tiles, int ntiles, const BLev *levs, const int *active,
                          int sel, bool fp32);

// The batched solver.
class BatchKMG {
  public:
    struct Problem {
        int nx, ny;
        double dx, dy;
        const mgp_geo *geo;         // host metrics, or
        const mgp_mapping *map;     // metrics generated on the GPU
        Bc bc;                      // types only (homogeneous)
    };
    BatchKMG(hipStream_t s, const std::vector<Problem> &probs, const mgp_solver_opts &o);
    ~BatchKMG();
    BatchKMG(const BatchKMG &) = delete;
    BatchKMG &operator=(const BatchKMG &) = delete;

    int count() const { return (int)P.size(); }
    const Grid &grid(int k) const { return P[k].g; }
    double *vec(int k, int which) { return P[k].v[which]; }
    // Solves every problem to ||rhs - L phi|| <= tol ||rhs||, from phi = 0 or
    // (zerophi == 0) from the resident phi.  iters[k] / outcome[k] per
    // problem; returns the V-cycle applications of the busiest problem.
    int solve(double tol, int maxiters, int zerophi, int *iters, int *outcome);
    int last_cycles = 0;

    // Step-wise use (the serving loop): start some problems while others run,
    // advance all running ones by one iteration, collect the finished ones.
    void start(const std::vector<int> &which, const std::vector<int> &warm, const std::vector<double> &tol,
               const std::vector<int> &maxit);
    std::vector<int> step();
    // step() in two halves, for callers that overlap several batches: the
    // launch, then (once step_done()) the checks; finish_step syncs only for
    // the true-residual checks it makes.
    void launch_step();
    bool step_done() const;
    std::vector<int> finish_step();
    int running() const;
    bool is_running(int k) const { return act_[k] != 0; }
    int iterations(int k) const { return it_[k]; }
    int outcome(int k) const { return oc_[k]; }
    // device bytes held by the batch (hierarchies, vectors, scalars, tables excluded)
    size_t footprint() const { return footprint_; }
    // the bytes one problem adds to a batch (its share of footprint())
    static size_t problem_bytes(const Problem &q, const mgp_solver_opts &o);
    // Measurement hook: mean ms of `reps` back-to-back launch sets of the
    // finest-level line relaxation of the even interior rows of every problem
    // (one launch per chunk class, side by side, as in a solve), timed with
    // HIP events on the batch's stream; *bytes the algorithmic bytes of one
    // set (each array element it needs once, the element size of the cycle).
    // Between solves only: it relaxes the resident finest-level corrections.
    double time_fine_xline(int reps, double *bytes);

    // Autotuner of the batched line solve's shape: the threads per row (64,
    // 128 or 256), and with them the chunk length, per key (nx, BC class, the
    // batch's precision).  The default is xline_batch_shape's rule.  The shape
    // decides how a row is partitioned into chunks, so a change moves the
    // solution by rounding; choices keyed this way (not timed per batch) keep
    // a problem's result the same alone and in any mix.
    struct TuneEntry {
        int nx, bc, fp32, threads;
    };
    static int bc_class(const Bc &bc);  // bit k: side k Dirichlet
    std::vector<TuneEntry> tuning() const;
    // Applies the entries whose key matches problems of the batch; returns the
    // number of problems whose shape changed.  Rebuilds tables and graphs.
    int set_tuning(const std::vector<TuneEntry> &t);
    // Picks the widths by timing the finest-level even-row line relaxation of
    // the whole batch (`reps` launch sets per candidate, HIP events), one band
    // of row lengths at a time; applies and returns the choice per key.
    // Between solves only: it relaxes the resident finest-level corrections.
    std::vector<TuneEntry> tune(int reps);

  private:
    void apply_shapes();
    double time_shapes(const std::vector<int> &C, const std::vector<int> &NT, int reps);
    void body();
    void precondition(int sel);
    void zebra(int d, int set, int sel, int k);
    // sc[slot] = sum of vol v[a] v[b] (and sc[slot2] of vol v[a] v[c], c >= 0)
    void reduce(int a, int b, int slot, const int *m, int c = -1, int slot2 = -1);
    void restart(const int *m);
    void true_residual(const int *m);
    void upload_flags(int *dst, const std::vector<int> &h);
    void fetch_scalars();

    hipStream_t s;
    int sd, su, sb;
    // per-problem iteration state
    std::vector<int> act_, it_, oc_, maxit_;
    std::vector<double> tol_, r0_, base_;
    size_t footprint_ = 0;
    std::vector<BProb> P;
    std::vector<BLev> L;
    std::vector<int> nlev;      // levels per problem
    int depth = 0;              // max levels
    void *arena = nullptr;
    double *scal = nullptr;     // all problems' BiCGStab scalars, SC_COUNT each
    BProb *dP = nullptr;
    BLev *dL = nullptr;
    int *active = nullptr, *mask = nullptr;
    // The x-line launches of one phase (one per chunk class) cover different
    // problems: they run side by side on streams forked from s and joined
    // back (in the captured iteration, parallel branches).
    std::vector<hipStream_t> aux;
    hipEvent_t fork_ev = nullptr;
    std::vector<hipEvent_t> join_ev;
    double *host_sc = nullptr;  // pinned: all problems' scalars
    // launch tables (device) with their lengths
    struct Table {
        void *p = nullptr;
        int n = 0;
        int C = 0;        // x-line tables: chunk length
        int threads = 0;  // x-line tables: threads per row
        size_t lds = 0;   // x-line tables: dynamic LDS
    };
    enum { SET_MID = 0, SET_BOTTOM = 1 };
    struct XlTable {
        int d, set, colour;
        Table t;
    };
    // the launch tables of a set of problems, with the captured iteration
    struct TableSet {
        Table t0;                    // finest-level tiles, item = problem
        Table tred;                  // reduction row blocks, item = problem (tx = block)
        std::vector<XlTable> xl;     // x-line task tables per (level, set, colour, chunk class)
        std::vector<Table> tr, tp;   // per level: restriction (coarse tiles) / prolongation (fine tiles)
        std::vector<int> bottom;     // per level: most coarsest-level sweeps of a problem ending there
        hipGraphExec_t graph = nullptr;
        int nprob = 0, depth = 0;
        // every table in one device buffer, staged in one pinned host buffer;
        // both are kept (and grown) across rebuilds, so that compacting the
        // set allocates nothing once it has been built at its largest
        void *dev = nullptr;
        unsigned char *host = nullptr;
        size_t dcap = 0, hcap = 0, used = 0;
    };
    Table stage(TableSet &ts, const void *host, size_t bytes, int n);
    void build_tables(TableSet &ts, const std::vector<int> &which);
    void release(TableSet &ts);       // the graph and the tables' contents (waits for the stream)
    void free_buffers(TableSet &ts);  // and the buffers
    TableSet full;               // every problem (built once)
    TableSet part;               // the problems still active, once they are at most half of `full`
    TableSet *T = &full;         // the set the launches use
    std::vector<int> Cp, NTp;    // chunk length and threads per row of each problem's line solves
    bool mixed = false;          // FP32 preconditioner cycle (method KRYLOV_MG32)
    bool use_graph = true;
};

// The serving loop over a batch of resident geometries ("slots"): requests
// (a slot and a rhs) queue up; each step admits queued requests to idle
// slots, advances every running solve by one iteration and retires the
// finished ones, whose solutions stay resident as the slot's warm start.
// Requests that arrive while others run join at the next step, so their
// work overlaps on the device; a request's result does not depend on what
// else is running (the batch keeps problems independent).
class Server {
  public:
    Server(const std::vector<BatchKMG::Problem> &slots, const mgp_solver_opts &o, double budget);
    ~Server();
    int slots() const { return (int)where.size(); }
    size_t footprint() const;
    // Admits a geometry at run time: its hierarchy is built (a batch of its
    // own) if the resident footprint stays within the budget, else
    // MGP_ERR_NOMEM and nothing changes.  Returns the new slot's index.
    int add_slot(const BatchKMG::Problem &q, const mgp_mapping *map);
    // Drops an idle slot (no request queued or running): its hierarchy and
    // warm start are freed (a batch shared with other slots is rebuilt
    // without it, keeping their warm starts).  Later submits to it fail.
    void evict(int slot);
    bool resident(int slot) const { return slot >= 0 && slot < slots() && where[slot].first >= 0; }
    long long submit(int slot, const double *rhs, double tol, int maxiters, int warm);
    // admits and advances; returns the solves on the device after the step
    int step();
    // done: 1 and the result (phi: nx x ny valid cells) if the request has
    // finished, then forgotten; 0 if it is queued or running
    int result(long long ticket, double *phi, int *iters, int *outcome);
    int queued() const { return (int)queue.size(); }

  private:
    struct Req {
        long long ticket;
        int slot, maxiters, warm;
        double tol;
        std::vector<double> rhs;
    };
    struct Res {
        std::vector<double> phi;
        int iters, outcome;
    };
    // Slots are grouped by size into classes, each a batch on a stream of
    // its own: a class advances at its own pace, and the classes' work
    // overlaps on the device.
    struct Class {
        hipStream_t s = nullptr;
        BatchKMG *B = nullptr;
        std::vector<int> slot;   // global slot of each local problem
        bool in_flight = false;
    };
    void admit(Class &c, int cls);
    void retire(Class &c, const std::vector<int> &finished);
    void free_class(Class &c);
    mgp_solver_opts opts_;
    double budget_ = 0.0;
    std::vector<BatchKMG::Problem> desc_;    // per slot: the geometry (map points into maps_)
    std::vector<std::unique_ptr<mgp_mapping>> maps_;
    std::vector<Class> cls;
    std::vector<std::pair<int, int>> where;  // per slot: (class, local index)
    std::vector<Req> queue;
    std::vector<long long> owner;  // per slot: running ticket, or 0
    std::vector<std::pair<long long, Res>> done;
    long long next = 1;
};

}  // namespace mgp
