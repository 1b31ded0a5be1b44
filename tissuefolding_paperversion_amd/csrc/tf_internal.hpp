// tf_internal.hpp -- device-resident state of a tf_handle and the launchers the C ABI drives.
//
// HBM layout (structure of arrays wherever a kernel walks elements or blocks in lanes):
//   element data   : conn[a*E + e] (int32), xref[k*E + e] (k = 3a+i), invfg/fg[q*E + e],
//                    lam/mu/eta[e], rate[c*E + e]
//   element output : kelem[e*324 + ab*9 + q] (36 node-pair blocks of 3x3, natural orientation),
//                    felem[e*18 + 3a + i] (negated elastic + internal viscous force)
//   node data      : uk/un/g/du [3N] interleaved (the reference's DoF order), sa[N], eta_ext[3N]
//   Jacobian (BSR3): row_ptr[N+1], col[nb], blk_row[nb], val[q*nb + blk] (component-major)
//   scatter maps   : blk_ptr[nb+1] / blk_src[36E] (contributing element blocks of every BSR
//                    block, sorted), node_ptr[N+1] / node_src[6E] (element corners of every node)
//   row tiles      : tile_r0/tile_r1/tile_seg -- runs of block rows of one tissue with at most
//                    tile_cap blocks and wg/3 rows, the unit of every row kernel and reduction
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/tf_nr.h"

#define TF_TILE_CAP_MAX 2048   // largest tile capacity (blocks); 48 KB of LDS products
#define TF_RED_SLOTS 16      // partial-sum slots per tile; partial[slot * T + tile]
#define TF_WARM_ITERS 4      // Newton iterations whose solves start from earlier timesteps' updates
#define TF_WARM_HIST 4       // updates kept per Newton iteration
#define TF_WARM_DOTS (TF_WARM_HIST * (TF_WARM_HIST + 1) / 2 + TF_WARM_HIST)  // Gram entries + right-hand side

// Per-tissue Krylov / NR scalars, one struct per segment, resident on the device.  Within a
// Krylov iteration the scalars (alpha, beta, omega, ...) are re-derived by every tile from the
// tiles' partial sums (see tf_krylov.hip).  The fields below are written by the first tile of a
// tissue only; a kernel never reads a field that it (or another tile of the same launch) writes,
// except the monotone converged/stalled flags, whose value every tile derives identically.
// Fields the single-kernel CG carries from launch to launch are indexed by launch parity P:
// launch P reads slot P and writes slot 1-P.
struct SegScalars {
    double bnorm2;           // ||b||^2
    double rnorm2;           // ||r||^2 of the iterate a stop (convergence, cap, breakdown) left
    double tol2;             // squared absolute stopping threshold
    double pad0;
    double alpha_hist[2];    // single-kernel CG: alpha of the previous iteration
    double gam_hist[2];      // single-kernel CG: gamma = (r, D^-1 r) of the previous iteration
    double best_rr_h[2];     // single-kernel CG: smallest ||r||^2 seen (stagnation test)
    int32_t iters_h[2];      // single-kernel CG: iterations done
    int32_t best_it_h[2];    // single-kernel CG: iteration at which best_rr_h was last halved
    int32_t converged;       // Krylov converged (or nothing to do)
    int32_t stalled;         // breakdown / iteration cap
    int32_t iters;           // Krylov iterations done (reported to the host)
    int32_t active;          // NR still iterating on this segment
};

// What the host reads once per Newton iteration (one copy): the assembly's status flags, the
// gate of the device-side update (ok) and why it stayed shut, then S SegScalars (the tissues'
// Krylov scalars after the solve's first round) and S doubles (||dU|| of every tissue).
struct NrReport {
    int32_t status[4];
    int32_t ok, unfinished, unusable, nan;
};
static_assert(sizeof(NrReport) == 32, "NrReport layout");
__host__ __device__ inline SegScalars* nr_report_seg(NrReport* r) { return reinterpret_cast<SegScalars*>(r + 1); }
__host__ __device__ inline double* nr_report_norm(NrReport* r, int S)
{
    return reinterpret_cast<double*>(nr_report_seg(r) + S);
}
inline size_t nr_report_bytes(int S) { return sizeof(NrReport) + (size_t)S * (sizeof(SegScalars) + sizeof(double)); }

// A linear solve that stopped short of its tolerance (iteration cap, stagnation, breakdown) with a
// residual at most 1e-2 of the right-hand side -- the loosest tolerance the inexact-Newton forcing
// asks for -- is a usable inexact Newton step.  Every stop records the squared residual of the
// iterate it leaves in rnorm2 (||b||^2 until the first stop test).  Host and device take the same
// decision from the same scalars.
__host__ __device__ inline bool solve_unusable(double rnorm2, double bnorm2) { return !(rnorm2 <= 1e-4 * bnorm2); }

// partial-sum slots
enum {
    SL_RHO0 = 0, SL_RHO1, SL_RR0, SL_RR1, SL_RV, SL_TS, SL_TT, SL_BB, SL_PQ, SL_RZ0, SL_RZ1, SL_NORM, SL_DL0, SL_DL1,
    SL_COUNT
};

struct KernelTimer {
    bool enabled = false;
    double total_ms = 0.0;
    int64_t launches = 0;
    std::vector<hipEvent_t> pending;  // start/stop pairs not yet read back
    // graph replays: events bracketing a whole replayed chunk
    hipEvent_t sample[2] = {nullptr, nullptr};
    double replay_ms = 0.0;           // whole replays
    int64_t replay_launches = 0;      // iterations launched by those replays (no-op launches included)
    int64_t replay_iters = 0;         // iterations the solves took
    bool capturing = false;           // inside a stream capture: no per-launch events
};

// Implicit packing: pairs of apical (or basal) nodes of one tissue closer than a threshold are
// pushed apart by a sigmoid force whose Jacobian enters M (blocks (a, b), (b, a) and the
// diagonal).  Per-tissue statistics stats[s * PK_NSTAT + k]:
#define PK_NSTAT 64
enum {
    PK_ST_T = 0,      // packingThreshold T (force scale)
    PK_ST_CELL = 1,   // spatial-hash cell edge (>= every detection threshold; 0: no detection)
    PK_ST_LO = 2,     // [3] minimum x, y, z of the tissue's nodes (hash origin)
    PK_ST_BB = 5,     // [4] reference bounding box: min x, min y, size x, size y
    PK_ST_GRID = 9,   // [50] detection threshold grid [10][5], row-major
    PK_ST_EXT = 59    // largest extent of the tissue's nodes over x, y, z
};
struct PackingState {
    // node and element tables (built at creation, detection tables on first use)
    std::vector<int8_t> h_node_pl;
    int8_t* node_pl = nullptr;        // [N] Node::tissuePlacement: 0 basal, 1 apical, 2 mid, 3 lateral
    uint8_t* elem_basal = nullptr;    // [E] 1: the element's basal side lengths count (basal or spanning)
    double* vref = nullptr;           // [E] reference volume
    double* mass = nullptr;           // [N] Node::mass (det(Fg) V_ref / 6 summed over the node's prisms)
    double* seg_stats = nullptr;      // [S][PK_NSTAT]
    int32_t *adj_ptr = nullptr, *adj_col = nullptr;  // element neighbours of every node, sorted
    int32_t* node_seg = nullptr;      // [N] tissue of every node
    int32_t* master_end = nullptr;    // [S] first node of the tissue's last detection chunk
    int4* chunks = nullptr;           // [n_chunks] runs of <= 256 elements: tissue, first element, end
    int32_t* seg_chunk_ptr = nullptr; // [S+1] runs of every tissue
    double* chunk_part = nullptr;     // [n_chunks][102] grid-bin partial sums of every run
    int32_t n_chunks = 0;
    int4* runs = nullptr;             // [n_runs] runs of <= 1024 nodes: tissue, first node, end
    int32_t* seg_run_ptr = nullptr;   // [S+1] runs of every tissue
    double* run_part = nullptr;       // [n_runs][10] bounding-box partials of every run
    int32_t n_runs = 0;
    double* ev = nullptr;             // [5][E] side-length values and centre of every element (scratch)
    double* vpn = nullptr;            // [E] det(Fg) V_ref / 6 of every element (scratch)
    int32_t H = 0;                    // hash slots (a power of two)
    int32_t *hcnt = nullptr, *hstart = nullptr, *cell_nodes = nullptr, *node_slot = nullptr;
    int32_t *pcount = nullptr, *pstart = nullptr, *scan_tmp = nullptr;
    // pair list (pacingNodeCouples0/1 and initialWeightPointx/y/z)
    int32_t n = 0, cap = 0;
    int32_t *a = nullptr, *b = nullptr;  // [cap] node ids
    double* w = nullptr;                 // [3 cap] weights
    int32_t* slot = nullptr;             // [2 cap] BSR blocks (a, b) and (b, a)
    std::vector<int32_t> h_a, h_b;
    int32_t nn = 0, node_cap = 0, ref_cap = 0;  // nodes with pairs; capacities of the node lists
    int32_t *nodes = nullptr, *nptr = nullptr, *nref = nullptr;  // their pairs: p (node a), ~p (node b)
    double *f = nullptr, *d = nullptr;   // [3N] packing force and Jacobian diagonal of every DoF
    std::vector<int64_t> h_blocks;       // pair blocks (min * N + max) kept in the BSR pattern
    hipEvent_t ev_count = nullptr;       // recorded behind the pair count's copy to the host
    hipGraphExec_t graph = nullptr;      // the detection's kernels up to the pair count, captured once
};

// Everything a Krylov tile kernel reads about its tile, in one 32-byte record: its tissue, BSR
// blocks [b0, b1), block rows [r0, r1) and the tissue's tiles [st0, st1).  The first 16 bytes hold
// what the first loads need (the tissue's state test, then the tile's blocks).
struct alignas(32) TileInfo {
    int32_t seg, b0, b1, r0, r1, st0, st1, pad;
};

struct tf_handle {
    // sizes
    int32_t N = 0, E = 0, nb = 0, S = 1, T = 0;
    int32_t wg = 256, tile_cap = 512;  // single-kernel CG workgroup size and tile capacity (blocks)
    int32_t max_lds = 65536;           // bytes of LDS one workgroup may allocate (device attribute)
    int32_t n_xcd = 1;                 // accelerator complex dies (XCDs) of the device
    int32_t n_cu = 256;                // compute units of the device
    int32_t nr_it = -1;                // Newton iteration of the running NR step (-1 outside one)
    // options
    tf_solver_opts opts;
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;

    // host copies needed for set-up and export
    std::vector<int32_t> h_row_ptr, h_col, h_node_seg_ptr, h_elem_seg_ptr, h_seg_tile_ptr;
    std::vector<int32_t> h_conn;  // [E*6]

    // element data
    int32_t *conn = nullptr;      // [6][E]
    double *xref = nullptr;       // [18][E]
    double *fmw = nullptr;        // [18][E] c_a of the mean deformation gradient Fm = sum_a x_a c_a^T
    double *fg = nullptr, *invfg = nullptr;  // [9][E]
    double *lam = nullptr, *mu = nullptr, *eta = nullptr;  // [E]
    double *rate = nullptr;       // [3][E]
    double *ginc = nullptr;       // [9][E] growth increments given by the host (replace the rates), or null
    uint8_t *exposed = nullptr;   // [E]
    int32_t *elem_seg = nullptr;  // [E]
    int any_viscous = 0;

    // node data
    double *uk = nullptr, *un = nullptr, *g = nullptr, *du = nullptr;  // [3N]
    double *eta_ext = nullptr;    // [3N]
    uint8_t *fixed = nullptr;     // [3N]
    double *sa = nullptr;         // [N]
    double *face_area = nullptr;  // [2][E]

    // assembly output / scratch
    double *kelem = nullptr;      // [E*324]
    double *felem = nullptr;      // [E*18]
    int32_t *status = nullptr;    // [4] device flags: 0 = non-positive det Fe count, 1 = rotated count,
                                  //   2 = packing partners past the list's room, 3 = non-finite det Fe count

    // Jacobian
    int32_t *row_ptr = nullptr, *col = nullptr, *blk_row = nullptr, *diag_blk = nullptr;
    double *val = nullptr;        // [9][nb]
    float *val32 = nullptr;       // [9][nb] (FP32 storage mode)
    double *dinv = nullptr;       // [N][9]
    int32_t *blk_ptr = nullptr, *blk_src = nullptr;
    int32_t *node_ptr = nullptr, *node_src = nullptr;
    std::vector<int32_t> h_node_ptr, h_node_src;

    // slave/master DoF binding (empty unless tf_set_bound_dofs was called) and ablated nodes
    std::vector<int32_t> h_bound;     // [2 * n_bound] (slave DoF, master DoF) in call order
    int32_t n_bound = 0, n_fold = 0, n_bind_nodes = 0, n_masters = 0;
    int32_t *fold_ptr = nullptr;      // [n_fold + 1] sources of each fold target
    int64_t *fold_tgt = nullptr;      // [n_fold] q*nb + blk of each target entry
    int64_t *fold_src = nullptr;      // q*nb + blk of each source entry (in val_base)
    double *fold_const = nullptr;     // [n_fold] constant term (1 on a slave diagonal)
    double *val_base = nullptr;       // [9][nb] M before the fold
    int32_t *bind_nodes = nullptr;    // [n_bind_nodes] nodes whose diagonal block the fold changes
    int32_t *bind_m_ptr = nullptr, *bind_m_dof = nullptr, *bind_s_dof = nullptr;  // masters -> slaves
    int32_t *bind_pairs = nullptr;    // [2 * n_bound]
    uint8_t *ablated = nullptr;       // [N] or null
    int32_t *blk_tr = nullptr;        // [nb] index of block (B, A) for block (A, B)

    // growth bookkeeping and discrete-adjoint work arrays (allocated on first use)
    double *rot_tz = nullptr;         // [E] growth-rotation angle applied by the last growth step (0: none)
    double *val_t = nullptr;          // [9][nb] transposed Jacobian values (transposed solves)
    double *g_save = nullptr;         // [3N]
    double *adj_xbar = nullptr, *adj_rhs = nullptr, *adj_pass = nullptr, *adj_y = nullptr, *adj_un_bar = nullptr;  // [3N]
    double *adj_invfg_bar = nullptr, *adj_fg_bar = nullptr, *adj_fg_prev = nullptr;  // [E][9]
    double *adj_rate_bar = nullptr;   // [E][3]
    double *adj_lam_bar = nullptr, *adj_mu_bar = nullptr, *adj_lam_acc = nullptr, *adj_mu_acc = nullptr, *adj_tz = nullptr;  // [E]

    // tiles and segments
    int32_t *tile_r0 = nullptr, *tile_r1 = nullptr, *tile_seg = nullptr;
    TileInfo* tile_info = nullptr;  // [T] the same per tile, with its block range and its tissue's tiles
    int32_t *seg_tile_ptr = nullptr;  // [S+1]
    int32_t *seg_node_ptr = nullptr;  // [S+1]
    int32_t *seg_counter = nullptr;   // [S] tickets for last-tile reductions
    double *partial = nullptr;        // [T][TF_RED_SLOTS]
    SegScalars *seg = nullptr;        // [S]
    SegScalars *h_seg = nullptr;      // pinned host mirror
    double *seg_norm = nullptr;       // [S] per-segment ||dU||
    double *h_seg_norm = nullptr;     // pinned
    int32_t *h_status = nullptr;      // pinned: [0, 4) copy of status, [4, 8) status after the growth
                                      // update, [8] the packing detection's pair count
    NrReport* nr_rep = nullptr;       // device report of a Newton iteration (nr_report_bytes(S))
    NrReport* hp_nr_rep = nullptr;    // pinned host copy
    double* hp_pos = nullptr;         // [3N] pinned, device-mapped staging of tf_get_positions
    double* dp_pos = nullptr;         //      its device address
    int32_t host_waits = 0;           // host waits for the device since the step began
    // Whole Newton iterations captured as graphs (nr_graph = 2), one per shape of their launches:
    // Newton iteration, kept warm-start updates, solver kind, value and tangent precision, first
    // burst, dt, forcing, ablated-node table.  Dropped with the Krylov chunks (pattern, tiling,
    // options or coarse space changed).
    struct NrGraphKey {
        int32_t it, m, kind, fp32, tangent, n_big, n_small, rtol_on;
        double dt;
        const void* ablated;
        bool operator==(const NrGraphKey& o) const
        {
            return it == o.it && m == o.m && kind == o.kind && fp32 == o.fp32 && tangent == o.tangent &&
                   n_big == o.n_big && n_small == o.n_small && rtol_on == o.rtol_on && dt == o.dt && ablated == o.ablated;
        }
    };
    struct NrGraph {
        NrGraphKey key;
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::vector<NrGraph> nr_graphs;
    uint64_t nr_graph_clock = 0;
    int32_t nr_graph_builds = 0;      // graphs captured since the step began
    int32_t persist_solves = 0;       // solves run as one persistent launch since the step began

    // Krylov work vectors [3N]
    double *kr = nullptr, *kr0 = nullptr, *kp = nullptr, *kph = nullptr, *kv = nullptr, *ks = nullptr,
           *ksh = nullptr, *kt = nullptr, *kx = nullptr, *km = nullptr, *kq = nullptr;

    // per-segment NR bookkeeping on the host
    std::vector<int32_t> seg_nr_iters, seg_converged;
    std::vector<int32_t> seg_failed;  // the tissue's linear solve stalled this step: it left the NR loop
    std::vector<double> seg_du;
    int32_t last_krylov_iters = 0;
    // inexact-Newton forcing, per tissue: the quadratic-convergence constant C estimated from
    // ||dU_k|| / ||dU_{k-1}||^2, the first update of the previous step, and this iteration's
    // Krylov relative tolerance (device copy read by the Krylov set-up while seg_rtol_on)
    std::vector<double> fc_C, fc_du0, fc_du_prev, h_seg_rtol2;
    double *seg_rtol2 = nullptr;      // [S]
    double* hp_rtol2 = nullptr;       // [S] pinned staging of h_seg_rtol2 (uploaded without a host wait)
    // the two-level prepare (column inverses, coarse refresh) already enqueued for the next solve
    bool tl_pre_valid = false, tl_pre_ok = false;
    bool seg_rtol_on = false;
    // Krylov warm start: for Newton iteration k < TF_WARM_ITERS, the updates that iteration took
    // in the last TF_WARM_HIST timesteps (warm_hist[k][0] the newest).  A solve starts from their
    // combination that minimises ||g - M x|| per tissue (warm_x0; warm_guess points at it during
    // the solve).  Work arrays: per-tile and per-tissue Gram entries; the coefficients are solved
    // on the device by every tile of the tissue.
    double* warm_hist[TF_WARM_ITERS][TF_WARM_HIST] = {};
    double* warm_block = nullptr;     // the buffers warm_hist points into (one allocation)
    int32_t warm_count[TF_WARM_ITERS] = {};
    double* warm_x0 = nullptr;
    double* warm_part = nullptr;      // [TF_WARM_DOTS][T]
    int32_t warm_part_T = 0;
    double* warm_seg = nullptr;       // [S][TF_WARM_DOTS]
    double** warm_tab = nullptr;      // [TF_WARM_ITERS][TF_WARM_HIST] device copy of warm_hist (kernel input)
    double** hp_warm_tab = nullptr;   // pinned staging of it
    const double* warm_guess = nullptr;
    int32_t fallbacks = 0;            // CG -> BiCGSTAB fallbacks since the last reset
    // Krylov iterations the solve of each Newton iteration took in the previous timestep: the
    // next solve of that iteration replays most of that many iterations before its first
    // convergence check (krylov_call: the Newton iteration being solved, -1 outside the loop)
    std::vector<int32_t> krylov_pred;
    int32_t krylov_call = -1;
    int32_t burst_slack = 0;          // small chunks a predicted first burst adds to the prediction
    bool no_persist = false;          // plan solves without the persistent launch (inside a graph capture)
    // A solve between its start (set-up and first round enqueued) and its end (tf_krylov.hip):
    // kind, chunk sizes, the rounds launched and the iterations they held.
    struct KrylovRun {
        bool pending = false, graph = false, predicted = false, timing = false, fp32 = false;
        bool persist = false;         // the two-level solve runs as one persistent launch (k_tp_run)
        bool persist_next = false;    // the next round_launch is that launch
        bool cg = false;              // opts.krylov is a CG kind (a stall may fall back to BiCGSTAB)
        int32_t kind = 0, chunk = 0, small = 0, n_big = 1, n_small = 0, round = 0, done_iters = 0, call = -1;
        int32_t last_launched = 0;    // iterations of the last round (its timing events are pending)
    } krun;
    bool tl_used = false;             // a solve of the current step ran the two-level PCG

    // graphs
    // one captured Krylov chunk per Jacobian value type (FP64, FP32)
    // captured Krylov chunks: [FP64 values, FP32 values] x [chunk, small chunk]
    hipGraphExec_t krylov_graph[4] = {nullptr, nullptr, nullptr, nullptr};
    int krylov_graph_iters[4] = {0, 0, 0, 0};
    int krylov_graph_kind[4] = {-1, -1, -1, -1};
    int krylov_graph_maxit[4] = {-1, -1, -1, -1};  // the iteration cap the captured launches carry

    // implicit packing
    PackingState pk;

    // two-level preconditioner (one tissue): block Jacobi plus a coarse space of the rigid-body
    // modes of node aggregates (columns through all layers, cut by an xy grid); see tf_krylov.hip
    // growth functions evaluated on the device (tf_growth.hip)
    struct GrowthFns {
        int n_fn = 0;
        std::vector<double> t0, t1;      // activity window of every function
        std::vector<int> kind;
        void* fns = nullptr;             // [n_fn] function records
        double* cells = nullptr;         // grid cells: corner growths, angles, eliminated flags
        double* rel = nullptr;           // [2][E] pinned relative positions
        double* z_init = nullptr;        // [E] initial relative heights
        int32_t* rel_src = nullptr;      // [E] element whose relative position an element takes
        double* box = nullptr;           // [S][PK_NSTAT] reference bounding boxes (PK_ST_BB)
        int pinned = 1, interpolation = 1;
        bool needs_box = false;          // a grid reads the current positions
    } gf;
    double time = 0.0;                   // simulation time at the start of the next timestep

    // Two-level preconditioner.  Every tissue has its own aggregates and coarse unknowns (tissue
    // after tissue); the coarse operator is block diagonal, one dense block per tissue.
    struct TwoLevel {
        bool built = false, fresh = false, failed = false;
        int32_t n_agg = 0, nc = 0, n_pairs = 0;
        int32_t age = 0;                                    // timesteps since the coarse operator was built
        std::vector<int32_t> base_it1;                      // [S] Krylov iterations of NR iteration 1 after the build
        std::vector<uint8_t> stale;                         // [S] a later NR iteration 1 took > 1.25x base_it1
        int32_t* agg_seg = nullptr;                         // [A] tissue of every aggregate
        int32_t* seg_c0 = nullptr;                          // [S+1] first coarse unknown of every tissue
        int32_t* seg_ldc = nullptr;                         // [S] row stride of the tissue's block of Ac
        int64_t* seg_acoff = nullptr;                       // [S] offset of that block in Ac
        uint8_t* refresh = nullptr;                         // [S] 1: the tissue's block is rebuilt
        std::vector<int32_t> h_seg_c0, h_seg_ldc, h_seg_lda;
        std::vector<int64_t> h_seg_acoff, h_seg_aoff;       // (Ainv offsets in halves)
        int32_t max_ldA = 0;                                // largest Ainv row stride (k_tp_z's LDS)
        int64_t ainv_halves = 0;                            // Ainv entries over all tissues
        int32_t* node_agg = nullptr;                        // [N] aggregate of every node
        int32_t *agg_ptr = nullptr, *agg_nodes = nullptr;   // [A+1], [N] nodes of every aggregate
        double* pcoef = nullptr;                            // [N][3][6] rows of P
        int32_t *coff = nullptr, *krank = nullptr;          // [A] coarse offset and rank
        int32_t *pair_a = nullptr, *pair_b = nullptr, *pair_ptr = nullptr, *pair_blk = nullptr;
        double* Ac = nullptr;                               // per tissue [ldc][ldc] P^T M P, then -Ac^-1 (lower tiles)
        int32_t ldc = 0;                                    // largest block stride: nc_s rounded up to the sweep's tile (64)
        double *swP = nullptr, *swO = nullptr, *swD = nullptr;  // the sweep's panels [ldc][64] and pivot inverses [2][64][64]
        void* Ainv = nullptr;                               // per tissue [nc_s][ldA_s] FP16 copy of Ac^-1 / scale
        double* ascale = nullptr;                           // [S][2] scale and its reciprocal
        int32_t ldA = 0;                                    // row stride of tissue 0's Ainv block (halves)
        int32_t L = 1, ncol = 0;                            // nodes per column, columns
        int32_t* col_blk = nullptr;                         // [ncol][L][L] BSR block of every node pair or -1
        float* dcinv = nullptr;                             // [ncol][3L][3L] FP32 inverse column blocks
        double* dcinv64 = nullptr;                          //   the same in FP64 (precond_fp64)
        double* Ainv64 = nullptr;                           // per tissue [nc_s][ldA_s] FP64 Ac^-1 (precond_fp64)
        int64_t* seg_aoff64 = nullptr;                      // [S] offset of the tissue's FP64 block
        int32_t* info = nullptr;                            // [S] the sweep's failure flag of every tissue
        // the solve in aggregate order (position r = node agg_nodes[r])
        int32_t *prow_ptr = nullptr, *pcol = nullptr, *pblk = nullptr;  // [N+1], [nb], [nb] permuted Jacobian pattern
        int32_t* pass_ptr = nullptr;                        // [A+1] row passes of every aggregate
        int2* pass_rows = nullptr;                          // [passes] rows [x, y) of every pass
        int4* wg_rec = nullptr;                             // [A][5] record of every workgroup of the aggregate-order
                                                            //     PCG (aggregates by descending size)
        double* pcoef_p = nullptr;                          // [N][3][6] rows of P by position
        double *val_p = nullptr;                            // [9][nb] FP64 values in aggregate order
        float* val32_p = nullptr;                           // [9][nb] FP32 values in aggregate order
        double *tpart = nullptr, *ptq = nullptr, *rcx = nullptr;  // [SL_COUNT][A], [nc], [2][nc]
        int32_t* tp_seg_ptr = nullptr;                      // [S+1] the aggregates as every tissue's "tiles"
        int32_t tp_max_blocks = 0;                          // most blocks of a pass (LDS products)
        // the persistent solve (one launch per solve): eligibility, the column slot of
        // every permuted block, every aggregate's neighbouring nodes, the grid-barrier words, and
        // how many workgroups of each kernel instance the device holds at once (-1: not queried)
        bool run_ok = false;
        int32_t *pslot = nullptr, *halo_ptr = nullptr, *halo_pos = nullptr;
        unsigned* bar = nullptr;
        int32_t run_cap[16] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
        std::vector<int32_t> h_node_agg;
    } tl;

    // timing
    KernelTimer t_elem, t_gather, t_spmv;
};

// ---- launchers (tf_assemble.hip) ------------------------------------------------------------
void launch_assemble_elements(tf_handle* h, double dt);
void launch_assemble_gather(tf_handle* h, double dt);
void launch_residual(tf_handle* h, double dt);
void launch_element_dump(tf_handle* h, int32_t e0, int32_t n, double dt, double* d_out);  // n prisms from e0
// Static scratch of the allocation-free element dump (tf_assemble.hip): room for
// TF_DUMP_SCRATCH_PRISMS prisms of TF_DUMP_DOUBLES values, one copy per GPU.  scratch_init resolves
// its device address for the current device (once per device); scratch_addr returns it (null
// before scratch_init ran on that device).
#define TF_DUMP_DOUBLES 684          // fe 18 + fv 18 + Ke 324 + Kv 324 per prism
#define TF_DUMP_SCRATCH_PRISMS 512   // 2.8 MB
#define TF_MAX_DEVICES 64
void scratch_init(int device);
double* scratch_addr(int device);
void launch_grow(tf_handle* h, double dt);
void launch_mean_f_weights(tf_handle* h);     // fmw from xref (once, after the reference upload)
void launch_surfaces(tf_handle* h);
void launch_copy(tf_handle* h, double* dst, const double* src, int64_t n);
void launch_fill(tf_handle* h, double* dst, double v, int64_t n);
void launch_seg_norm(tf_handle* h, const double* x);   // per-segment ||x||_2 into h->seg_norm
void launch_update_uk(tf_handle* h, const NrReport* gate = nullptr);  // uk += du on active segments (if gate->ok)
void build_tiles(tf_handle* h, int wg, int tile_cap);  // (re)partitions rows into tiles (tf_api.hip)
bool tiling_fits(const tf_handle* h, int wg, int tile_cap);  // LDS of the single-kernel CG fits
void drop_krylov_graphs(tf_handle* h);  // destroys the captured Krylov chunks

// The precision of the current Newton iteration: FP64 tangents and Jacobian values from
// opts.mixed_fp64_from on, the requested reduced precision before it and outside the NR loop.
inline bool nr_late(const tf_handle* h)
{
    return h->opts.mixed_fp64_from >= 0 && h->nr_it >= h->opts.mixed_fp64_from;
}
inline int tangent_precision_now(const tf_handle* h) { return nr_late(h) ? TF_PREC_FP64 : h->opts.tangent_precision; }
inline bool fp32_values(const tf_handle* h) { return h->opts.matrix_precision == TF_PREC_FP32 && !nr_late(h); }
void build_pattern(tf_handle* h);  // BSR pattern, scatter lists and bound-DoF fold lists (tf_api.hip)
void launch_bind_matrix(tf_handle* h);    // M <- N^T M N + I_slaves on the fold targets
void launch_bind_residual(tf_handle* h);  // g <- N^T g
void launch_bind_du(tf_handle* h);        // du[slave] <- du[master]
// ---- implicit packing (tf_packing.hip) -------------------------------------------------------
void packing_init(tf_handle* h, const int8_t* node_placement, const uint8_t* exposed, const double* ref_pos_aos);
void packing_free(tf_handle* h);
void launch_reference_box(tf_handle* h, double* stats);  // PK_ST_LO/BB/EXT of every tissue at uk
// growth functions (tf_growth.hip)
#define TF_GROWTH_MAX_ACTIVE 16
void growth_functions_set(tf_handle* h, int n, const tf_growth_function* f, const double* z_init,
                          const int32_t* rel_src, int pinned, int interpolation);
void growth_functions_eval(tf_handle* h, double dt);  // increments of the active functions into ginc
void growth_functions_pin(tf_handle* h);
void growth_functions_free(tf_handle* h);
void packing_prepare(tf_handle* h);             // detection tables (first use)
void packing_reserve(tf_handle* h, int n);      // pair arrays for n pairs
int packing_detect(tf_handle* h);               // masses + detection on uk; returns the pair count
void packing_detect_launch(tf_handle* h);       // the same, enqueued up to the pair count's copy
int packing_detect_finish(tf_handle* h);        // waits for the count (not for later work), installs
void packing_install(tf_handle* h, const std::vector<int32_t>& a, const std::vector<int32_t>& b);
void packing_clear(tf_handle* h);
void launch_packing_mass(tf_handle* h);
void launch_packing_nodes(tf_handle* h);        // pk.f, pk.d of the nodes with pairs at uk
void launch_packing_offdiag(tf_handle* h);      // pair blocks of M, after k_gather
// ---- discrete adjoint (tf_adjoint.hip) -----------------------------------------------------
void launch_adj_elements(tf_handle* h, const double* y, double* invfg_bar, double* lam_bar, double* mu_bar);
void launch_adj_un(tf_handle* h, double dt, const double* y, double* un_bar);
void launch_adj_growth(tf_handle* h, double dt, const double* theta_z, const double* fg_prev, const double* invfg_bar,
                       double* fg_bar, double* rate_bar);
void launch_spmv_t(tf_handle* h, const double* x, double* y);
void launch_transpose_vals(tf_handle* h, double* val_t);
void launch_dinv_all(tf_handle* h);
void launch_adj_split(tf_handle* h, const double* xbar, double* rhs, double* pass);
void launch_accumulate(tf_handle* h, const double* a, double* b, int64_t n);
// ---- Krylov (tf_krylov.hip) -----------------------------------------------------------------
int krylov_solve(tf_handle* h);  // solves M du = g; returns tf_status
// The same solve in two halves: krylov_solve_begin enqueues the set-up and the first round
// without a host wait and returns true when that round is a predicted burst (a solve that may be
// complete without a check); krylov_solve_end finishes it (seg_known: h_seg already holds the
// tissues' scalars after that round; scattered: du already holds the solution) with the BiCGSTAB
// fallback and the copy of the solution into du.
bool krylov_solve_begin(tf_handle* h);
int krylov_solve_end(tf_handle* h, bool seg_known, bool scattered);
// krylov_solve_begin in two halves: the host's plan of the solve (kind, first round; no launch) and
// its launches (direct: the first round's iterations enqueued as kernels, for a capture)
bool krylov_solve_plan(tf_handle* h);
void krylov_solve_launch(tf_handle* h, bool direct);
void launch_solution_to_du(tf_handle* h);  // du <- the running solve's x (node order)
// Warm start of the solve of Newton iteration k: x0 in warm_x0 from the kept updates (zero on a
// tissue whose least-squares problem is degenerate); false when no update is kept.
bool warm_start_begin(tf_handle* h, int k, bool launch = true);  // Gram, coefficients, x0 on the device, no wait
void warm_start_launch(tf_handle* h, int k);  // the kernels of warm_start_begin (buffers allocated by it)
void krylov_pre_solve(tf_handle* h, bool launch = true);  // the next solve's two-level prepare
void krylov_pre_launch(tf_handle* h);         // its per-solve kernels (two-level solves), when krylov_pre_solve
                                              // did not launch them
bool krylov_pre_solve_rebuilds(const tf_handle* h);  // krylov_pre_solve would build or rebuild the coarse space
void warm_start_push(tf_handle* h, int k);    // keeps du as the newest update of iteration k
void warm_start_free(tf_handle* h);
void warm_start_clear(tf_handle* h);
void twolevel_free(tf_handle* h);
void twolevel_invalidate(tf_handle* h);  // the pattern changed
void twolevel_setup(tf_handle* h);       // builds the coarse space if twolevel_active and not built
int twolevel_export(tf_handle* h, int32_t* nc, double* Ac, double* Ainv);  // -1: no fresh coarse operator
// ms per Krylov iteration (whole replays) of the solver krylov_solve would run for a symmetric
// system: the two-level PCG when twolevel_active, else the single-kernel CG
double krylov_time_iteration(tf_handle* h, int n_iters);
bool twolevel_active(const tf_handle* h);  // the CG solves of this handle use the two-level preconditioner
// Average device time (ms) of one launch of the dominant Krylov kernel (the one tf_kernel_bytes
// prices under "spmv"), n_launches replayed back to back on the current Jacobian; 0 for BiCGSTAB
// which: 0 the SpMV kernel, 1 k_tp_z, 2 ms per iteration of the persistent two-level solve (one launch)
double krylov_probe(tf_handle* h, int n_launches, int which = 0);
#define TF_TWO_LEVEL_MIN_NODES 4096       // TF_PRECOND_AUTO: two-level from this many nodes
#ifndef TF_ONE_READ_SLACK
#define TF_ONE_READ_SLACK 1               // small chunks a one-read iteration's first burst adds to its prediction
#endif
#define TF_GRAPH_NODES 32                 // kernel launches per captured Krylov graph (at most)
void launch_spmv(tf_handle* h, const double* x, double* y);

// error helpers
#define TF_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) throw tf_hip_error(e_, #call, __FILE__, __LINE__);     \
    } while (0)

struct tf_hip_error {
    hipError_t code;
    std::string what;
    tf_hip_error(hipError_t c, const char* call, const char* file, int line) : code(c) {
        what = std::string(hipGetErrorString(c)) + " in " + call + " at " + file + ":" + std::to_string(line);
    }
};

// timing helpers (record events around a launch on the handle's stream)
void timer_begin(tf_handle* h, KernelTimer& t, hipEvent_t* ev);
void timer_end(tf_handle* h, KernelTimer& t, hipEvent_t* ev);
void timer_collect(KernelTimer& t);
