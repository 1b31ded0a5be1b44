// tf_packing.hip -- implicit packing: node masses, per-tissue pair detection on a spatial hash,
// and the pairs' sigmoid forces and Jacobian entries (gfx950, FP64).
//
// Detection (once per timestep, SC/Simulation.cpp:8167-8440) runs in five steps:
//   k_pk_vpn, k_pk_mass : det(Fg) V_ref / 6 per prism, summed per node (updateNodeMasses).
//   k_pk_evals, k_pk_box(_finish), k_pk_bins, k_pk_finish : side-length averages and centres per
//                prism; the reference's bounding box per tissue from partials over runs of 1024
//                nodes (its running-minimum rule included);
//                grid-bin partial sums per run of 256 prisms; per tissue, the runs' partials in
//                order into the 10 x 5 threshold grid and the packing threshold T.
//   k_pk_cells : eligible nodes (mass > 0, apical or basal) go to a hashed uniform grid whose
//                cell is at least the largest detection threshold (Teschner et al., "Optimized
//                spatial hashing for collision detection of deformable objects", VMV 2003), as
//                a counting sort: count, exclusive scan, fill.
//   k_pk_count, k_pk_fill : every master node's 27 neighbour cells are searched for the slaves
//                j > i of its tissue that the reference's O(N^2) scan would accept; the count
//                pass (32 lanes per master) sizes the list, the fill pass (only when pairs exist)
//                writes every master's pairs in ascending slave order, so that the list comes
//                out in the reference's order (master, then slave).
// Per NR iteration (SC/Simulation.cpp:6489-6499, :7149-7295):
//   k_pk_nodes   : one lane per node with pairs sums its pairs' forces and Jacobian diagonal in
//                  pair order (consumed by k_residual and k_gather).
//   k_pk_offdiag : one lane per pair adds its off-diagonal entries to blocks (a, b) and (b, a),
//                  which the BSR pattern carries for every pair.
#include "tf_internal.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace {

constexpr double PK_MULT = 5000.0;  // packingMultiplier (SC/Simulation.cpp:8187)
constexpr double PK_SAT = 5.0;      // sigmoidSaturationForPacking (SC/Simulation.cpp:8188)
constexpr int PK_MAX_PARTNERS = 64; // slaves one master may collect in one detection
constexpr int PK_CHUNK = 256;       // elements per workgroup of the grid-bin kernel
constexpr int PK_NODE_RUN = 1024;   // nodes per workgroup of the bounding-box kernel
constexpr int PK_NPART = 102;       // partials per element run: 50 bin sums, 50 bin counts, 2 side-length sums
constexpr int SCAN_WG = 1024, SCAN_PER = 4, SCAN_TILE = SCAN_WG * SCAN_PER;

template <typename T>
T* pk_alloc(size_t n)
{
    if (n == 0) n = 1;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) throw tf_hip_error(e, "hipMalloc", __FILE__, __LINE__);
    return (T*)p;
}
template <typename T>
void pk_free(T*& p)
{
    if (p) { (void)hipFree(p); p = nullptr; }
}
template <typename T>
void pk_h2d(tf_handle* h, T* dst, const T* src, size_t n)
{
    if (!n) return;
    TF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
    TF_HIP(hipStreamSynchronize(h->stream));
}

// Grid bin of a relative coordinate: floor(v) clamped to [0, n-1] (negative and NaN -> 0).
__device__ __forceinline__ int bin_index(double v, int n)
{
    if (!(v >= 0.0)) return 0;
    if (v >= (double)n) return n - 1;
    return (int)v;
}

__device__ __forceinline__ uint32_t cell_hash(int s, int cx, int cy, int cz, uint32_t mask)
{
    return ((uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u ^ (uint32_t)cz * 83492791u ^ (uint32_t)s * 2654435761u) &
           mask;
}

// Prism::getApicalSideLengthAverage / getBasalSideLengthAverage (SC/Prism.cpp:1443-1490).
__device__ double side_average(const double* __restrict__ uk, const int c[6], int apical)
{
    const int o = apical ? 3 : 0;
    const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    double dsum = 0.0;
    int counter = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double* p = uk + 3 * c[o + pr[k][0]];
        const double* q = uk + 3 * c[o + pr[k][1]];
        const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > 0) {
            dsum += sqrt(d2);
            counter++;
        }
    }
    if (counter > 0) dsum /= counter;
    return dsum;
}


// ---- node masses (SC/Simulation.cpp:7552-7560, SC/ShapeBase.cpp:1243-1245, :3938-3942) -----
// VolumePerNode = det(Fg) V_ref / 6 per prism, then every node sums its prisms' values in
// element order.
__global__ __launch_bounds__(256) void k_pk_vpn(const double* __restrict__ fg, const double* __restrict__ vref, int E,
                                                double* __restrict__ vpn)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double F[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) F[q] = fg[(size_t)q * E + e];
    const double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                       F[2] * (F[3] * F[7] - F[4] * F[6]);
    vpn[e] = det * vref[e] / 6.0;
}

__global__ __launch_bounds__(256) void k_pk_mass(const int32_t* __restrict__ node_ptr, const int32_t* __restrict__ node_src,
                                                 const double* __restrict__ vpn, const uint8_t* __restrict__ ablated, int N,
                                                 double* __restrict__ mass)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double m = 0.0;
    for (int k = node_ptr[n]; k < node_ptr[n + 1]; ++k) m += vpn[node_src[k] / 6];
    mass[n] = (ablated && ablated[n]) ? 0.0 : m;
}

// ---- per-element side lengths and centres (one lane per element) ---------------------------
// ev[0]: the element's contribution to the columnar side-length sum (its positive apical
// average, plus its positive basal average when basal sides count), ev[1]: how many averages
// that is, ev[2]: the value its grid bin sums (the basal average when basal sides count, else
// the apical one), ev[3], ev[4]: centre x, y (mean of the six nodes, ShapeBase::getCentre,
// SC/ShapeBase.cpp:240-250).  Structure of arrays: ev[k * E + e].
__global__ __launch_bounds__(256) void k_pk_evals(const double* __restrict__ uk, const int32_t* __restrict__ conn, int E,
                                                  const uint8_t* __restrict__ elem_basal, double* __restrict__ ev)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int c[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) c[a] = conn[(size_t)a * E + e];
    double cur = side_average(uk, c, 1);
    double sum = 0.0, cnt = 0.0;
    if (cur > 0) { sum += cur; cnt += 1.0; }
    if (elem_basal[e]) {
        cur = side_average(uk, c, 0);
        if (cur > 0) { sum += cur; cnt += 1.0; }
    }
    double cx = 0.0, cy = 0.0;
#pragma unroll
    for (int a = 0; a < 6; ++a) { cx += uk[3 * c[a]]; cy += uk[3 * c[a] + 1]; }
    ev[e] = sum;
    ev[(size_t)E + e] = cnt;
    ev[2 * (size_t)E + e] = cur;
    ev[3 * (size_t)E + e] = cx / 6;
    ev[4 * (size_t)E + e] = cy / 6;
}

// ---- per-tissue bounding box, thresholds and grid (one workgroup per tissue) -----------------
// Simulation::calculateBoundingBox (SC/Simulation.cpp:9148-9189), relative positions
// (SC/ShapeBase.cpp:1842-1848), Simulation::getAverageSideLength (SC/Simulation.cpp:3205-3265)
// and the 0.4 scalings of detectPacingNodes (:8175-8186).
// The box: "if (x < min) min = x; else if (x > max) max = x" from (1e5, -1e5) leaves out of
// the maximum every node that lowers the running minimum.  Such a node lies below all earlier
// ones, so it cannot be the largest of the nodes after the initial strictly decreasing run
// x_0 > x_1 > ... > x_{r-1}; every node of that run lowers the minimum.  Hence
// max = max_{i >= r} x_i with r the first i >= 1 where x_i >= x_{i-1}.
// The grid: bins sum the element's bin value; their counters are shared between bins (i, 4)
// and (i + 1, 0), as the reference's 10 x 4 counter array indexed with j < 5 does.
// One workgroup per run of PK_NODE_RUN nodes of one tissue (4 per lane): the run's minima and
// maxima of x, y, z, its maxima of x and y, and its first node n > n0 with x_n >= x_{n-1} (y
// likewise; the tissue's end n1 if none).  part[c][k]: 0-2 minima, 3-5 maxima, 6-7 x and y maxima,
// 8-9 the first nodes.  Minima and maxima are exact in any order.
__global__ __launch_bounds__(256) void k_pk_box(const int4* __restrict__ runs, const int32_t* __restrict__ seg_node_ptr,
                                                const double* __restrict__ uk, double* __restrict__ part)
{
    __shared__ double sh[4][10];
    const int4 rn = runs[blockIdx.x];  // tissue, first node, end
    const int n0 = seg_node_ptr[rn.x], n1 = seg_node_ptr[rn.x + 1];
    const int t = threadIdx.x, lane = t % 64, w = t / 64;
    double v[8] = {1.0e300, 1.0e300, 1.0e300, -1.0e300, -1.0e300, -1.0e300, 0.0, 0.0};
    int rx = n1, ry = n1;  // none in this run: the tissue's end
#pragma unroll
    for (int k = 0; k < PK_NODE_RUN / 256; ++k) {
        const int n = rn.y + t + 256 * k;
        if (n < rn.z) {
            const double x = uk[3 * n], y = uk[3 * n + 1], z = uk[3 * n + 2];
            v[0] = fmin(v[0], x); v[1] = fmin(v[1], y); v[2] = fmin(v[2], z);
            v[3] = fmax(v[3], x); v[4] = fmax(v[4], y); v[5] = fmax(v[5], z);
            if (n > n0) {
                if (x >= uk[3 * (n - 1)]) rx = min(rx, n);
                if (y >= uk[3 * (n - 1) + 1]) ry = min(ry, n);
            }
        }
    }
    double fx = (double)rx, fy = (double)ry;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], __shfl_xor(v[d], off, 64));
            v[3 + d] = fmax(v[3 + d], __shfl_xor(v[3 + d], off, 64));
        }
        fx = fmin(fx, __shfl_xor(fx, off, 64));
        fy = fmin(fy, __shfl_xor(fy, off, 64));
    }
    if (lane == 0) {
#pragma unroll
        for (int d = 0; d < 6; ++d) sh[w][d] = v[d];
        sh[w][8] = fx;
        sh[w][9] = fy;
    }
    __syncthreads();
    if (t < 10 && t != 6 && t != 7) {
        double r = sh[0][t];
        for (int k = 1; k < 4; ++k) r = (t < 3 || t >= 8) ? fmin(r, sh[k][t]) : fmax(r, sh[k][t]);
        part[(size_t)blockIdx.x * 10 + t] = r;
        if (t == 3 || t == 4) part[(size_t)blockIdx.x * 10 + t + 3] = r;  // the x and y maxima
    }
}

// One workgroup per tissue: the runs' partials into the reference's bounding box.  The box's
// maximum x is the largest x over the nodes from the first n > n0 with x_n >= x_{n-1} on (the
// nodes before it form the strictly decreasing start that only lowers the running minimum):
// whole runs past that node give their maxima, the run holding it is read again from it on.
// Minima and maxima are exact in any order, so the lanes reduce them together: a butterfly in
// every wave, then the waves' results.
__global__ __launch_bounds__(256) void k_pk_box_finish(const int32_t* __restrict__ seg_run_ptr,
                                                       const int32_t* __restrict__ seg_node_ptr,
                                                       const int4* __restrict__ runs, const double* __restrict__ part,
                                                       const double* __restrict__ uk, double* __restrict__ stats)
{
    __shared__ double sh[4][8];
    const int s = blockIdx.x, t = threadIdx.x, lane = t % 64, w = t / 64;
    const int c0 = seg_run_ptr[s], c1 = seg_run_ptr[s + 1];
    const int n0 = seg_node_ptr[s], n1 = seg_node_ptr[s + 1];
    // v: minima x, y, z, maxima x, y, z, first nodes x, y (exact as doubles)
    double v[8] = {1.0e300, 1.0e300, 1.0e300, -1.0e300, -1.0e300, -1.0e300, (double)n1, (double)n1};
    for (int c = c0 + t; c < c1; c += 256) {
        const double* pc = part + (size_t)c * 10;
#pragma unroll
        for (int d = 0; d < 3; ++d) { v[d] = fmin(v[d], pc[d]); v[3 + d] = fmax(v[3 + d], pc[3 + d]); }
        v[6] = fmin(v[6], pc[8]);
        v[7] = fmin(v[7], pc[9]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double o = __shfl_xor(v[k], off, 64);
            v[k] = (k < 3 || k >= 6) ? fmin(v[k], o) : fmax(v[k], o);
        }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 8; ++k) sh[w][k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double r = sh[0][k];
        for (int j = 1; j < 4; ++j) r = (k < 3 || k >= 6) ? fmin(r, sh[j][k]) : fmax(r, sh[j][k]);
        v[k] = r;
    }
    const int rx = (int)v[6], ry = (int)v[7];
    double qx = -100000.0, qy = -100000.0;
    for (int c = c0 + t; c < c1; c += 256) {
        const int4 rn = runs[c];
        if (rn.y >= rx) qx = fmax(qx, part[(size_t)c * 10 + 6]);
        if (rn.y >= ry) qy = fmax(qy, part[(size_t)c * 10 + 7]);
    }
    // the runs holding rx and ry (runs start at n0 + k PK_NODE_RUN): from that node to the run's end
    if (rx < n1) {
        const int end = min(n0 + ((rx - n0) / PK_NODE_RUN + 1) * PK_NODE_RUN, n1);
        for (int n = rx + t; n < end; n += 256) qx = fmax(qx, uk[3 * n]);
    }
    if (ry < n1) {
        const int end = min(n0 + ((ry - n0) / PK_NODE_RUN + 1) * PK_NODE_RUN, n1);
        for (int n = ry + t; n < end; n += 256) qy = fmax(qy, uk[3 * n + 1]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        qx = fmax(qx, __shfl_xor(qx, off, 64));
        qy = fmax(qy, __shfl_xor(qy, off, 64));
    }
    __syncthreads();  // sh is reused
    if (lane == 0) { sh[w][0] = qx; sh[w][1] = qy; }
    __syncthreads();
    if (t == 0) {
        for (int j = 1; j < 4; ++j) { qx = fmax(qx, sh[j][0]); qy = fmax(qy, sh[j][1]); }
        const double bx = fmin(100000.0, v[0]), by = fmin(100000.0, v[1]);
        double* st = stats + (size_t)s * PK_NSTAT;
        st[PK_ST_LO + 0] = v[0];
        st[PK_ST_LO + 1] = v[1];
        st[PK_ST_LO + 2] = v[2];
        st[PK_ST_BB + 0] = bx;
        st[PK_ST_BB + 1] = by;
        st[PK_ST_BB + 2] = qx - bx;
        st[PK_ST_BB + 3] = qy - by;
        st[PK_ST_EXT] = fmax(v[3] - v[0], fmax(v[4] - v[1], v[5] - v[2]));
    }
}

// One workgroup per run of PK_CHUNK elements of one tissue: every element's grid bin, and the
// run's per-bin sums of the bin value and element counts, plus the side-length sum and count.
// Inside a wave the lanes of one bin are summed by a butterfly, bin after bin (a wave of
// neighbouring elements holds few bins); the waves' results are added in wave order, so every
// partial has a fixed order.  part[c][k]: k < 50 bin sums, 50 <= k < 100 bin counts, 100 the
// side-length sum, 101 its count.
__global__ __launch_bounds__(PK_CHUNK) void k_pk_bins(const int4* __restrict__ chunks, const double* __restrict__ ev, int E,
                                                      const double* __restrict__ stats, double* __restrict__ part)
{
    constexpr int W = PK_CHUNK / 64;
    __shared__ double wsum[W][50], wcnt[W][50], wd[W][2];
    const int4 ck = chunks[blockIdx.x];  // tissue, first element, end
    const int t = threadIdx.x, lane = t % 64, w = t / 64;
    for (int k = t; k < W * 50; k += PK_CHUNK) { (&wsum[0][0])[k] = 0.0; (&wcnt[0][0])[k] = 0.0; }
    const double* st = stats + (size_t)ck.x * PK_NSTAT;
    const int e = ck.y + t;
    const bool own = e < ck.z;
    int b = -1;
    double v = 0.0, d0 = 0.0, d1 = 0.0;
    if (own) {
        d0 = ev[e];
        d1 = ev[(size_t)E + e];
        v = ev[2 * (size_t)E + e];
        const int bx = bin_index((ev[3 * (size_t)E + e] - st[PK_ST_BB]) / st[PK_ST_BB + 2], 10);
        const int by = bin_index(((ev[4 * (size_t)E + e] - st[PK_ST_BB + 1]) / st[PK_ST_BB + 3]) / 2.0, 5);
        b = 5 * bx + by;
    }
    __syncthreads();
    unsigned long long pending = __ballot(own);
    while (pending) {  // wave-uniform
        const int bl = __shfl(b, __ffsll(pending) - 1, 64);
        const unsigned long long m = __ballot(b == bl);
        double x = b == bl ? v : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) { wsum[w][bl] = x; wcnt[w][bl] = (double)__popcll(m); }
        pending &= ~m;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        d0 += __shfl_xor(d0, off, 64);
        d1 += __shfl_xor(d1, off, 64);
    }
    if (lane == 0) { wd[w][0] = d0; wd[w][1] = d1; }
    __syncthreads();
    if (t < PK_NPART) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < W; ++k) acc += t < 50 ? wsum[k][t] : t < 100 ? wcnt[k][t - 50] : wd[k][t - 100];
        part[(size_t)blockIdx.x * PK_NPART + t] = acc;
    }
}

// One workgroup per tissue: the runs' partials (lane group j takes runs j, j + 10, ... in order,
// then the 10 group sums are added in group order), then the 10 x 5 grid, the threshold T and
// the hash cell.
constexpr int PK_FIN_G = 10;  // lane groups of PK_NPART lanes (1020 of 1024 lanes)
__global__ __launch_bounds__(1024) void k_pk_finish(const int32_t* __restrict__ seg_chunk_ptr,
                                                    const double* __restrict__ part, double* __restrict__ stats)
{
    __shared__ double grp[PK_FIN_G][PK_NPART];
    __shared__ double tot[PK_NPART];
    const int s = blockIdx.x, t = threadIdx.x;
    const int c0 = seg_chunk_ptr[s], c1 = seg_chunk_ptr[s + 1];
    const int j = t / PK_NPART, k = t % PK_NPART;
    if (j < PK_FIN_G) {
        double acc = 0.0;
#pragma unroll 4
        for (int c = c0 + j; c < c1; c += PK_FIN_G) acc += part[(size_t)c * PK_NPART + k];
        grp[j][k] = acc;
    }
    __syncthreads();
    if (t < PK_NPART) {
        double acc = grp[0][t];
        for (int g = 1; g < PK_FIN_G; ++g) acc += grp[g][t];
        tot[t] = acc;
    }
    __syncthreads();
    if (t >= 64) return;
    double* st = stats + (size_t)s * PK_NSTAT;
    double g = 0.0;
    if (t < 50) {
        const double sb = tot[t], cb = tot[50 + t];
        if (t == 0) {  // alias index 0 belongs to bin (0, 0) alone
            g = 0.4 * (cb > 0 ? sb / cb : sb);
        } else if (cb > 0) {  // other bins: only with a box older than the positions
            const int alias = 4 * (t / 5) + t % 5;
            double ca = 0.0;
            for (int b = 0; b < 50; ++b)
                if (4 * (b / 5) + b % 5 == alias) ca += tot[50 + b];
            g = 0.4 * (ca > 0 ? sb / ca : sb);
        }
        st[PK_ST_GRID + t] = g;
    }
    double hc = g;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) hc = fmax(hc, __shfl_xor(hc, off, 64));
    if (t == 0) {
        const double ds = tot[100], dc = tot[101];
        st[PK_ST_T] = dc > 0 ? 0.4 * (ds / dc) : 0.0;
        st[PK_ST_CELL] = hc > 0 ? fmax(hc, st[PK_ST_EXT] / 1048576.0) : 0.0;
    }
}

// ---- spatial hash ------------------------------------------------------------------------
__device__ __forceinline__ bool can_pack(const double* __restrict__ mass, const int8_t* __restrict__ pl, int n)
{
    // Node::checkIfNodeHasPacking (SC/Node.cpp:133-151); no lateral elements in this model
    return mass[n] != 0.0 && (pl[n] == 0 || pl[n] == 1);
}

__global__ __launch_bounds__(256) void k_pk_cells(int N, const int32_t* __restrict__ node_seg,
                                                  const double* __restrict__ stats, const double* __restrict__ uk,
                                                  const double* __restrict__ mass, const int8_t* __restrict__ pl,
                                                  uint32_t hmask, int32_t* __restrict__ node_slot, int32_t* __restrict__ hcnt)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int s = node_seg[n];
    const double* st = stats + (size_t)s * PK_NSTAT;
    const double hc = st[PK_ST_CELL];
    if (!(hc > 0.0) || !can_pack(mass, pl, n)) { node_slot[n] = -1; return; }
    const int cx = (int)((uk[3 * n] - st[PK_ST_LO]) / hc);
    const int cy = (int)((uk[3 * n + 1] - st[PK_ST_LO + 1]) / hc);
    const int cz = (int)((uk[3 * n + 2] - st[PK_ST_LO + 2]) / hc);
    const uint32_t slot = cell_hash(s, cx, cy, cz, hmask);
    node_slot[n] = (int32_t)slot;
    atomicAdd(hcnt + slot, 1);
}

__global__ __launch_bounds__(256) void k_pk_fill_cells(int N, const int32_t* __restrict__ node_slot,
                                                       int32_t* __restrict__ cursor, int32_t* __restrict__ cell_nodes)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int slot = node_slot[n];
    if (slot >= 0) cell_nodes[atomicAdd(cursor + slot, 1)] = n;
}

// ---- pair search ----------------------------------------------------------------------------
// The acceptance test of detectPacingNodes (SC/Simulation.cpp:8238-8299) for master i and slave
// j > i of the same tissue: same placement, not sharing a prism (Node::immediateNeigs,
// SC/ShapeBase.cpp:3105-3122), |x_i - x_j|^2 < t^2 with t the grid threshold of the master's bin;
// masters from the tissue's last detection chunk are dropped as the reference's merge loop does.
struct Master {
    double x, y, z, t2;
    int cx, cy, cz, s, nend, a0, a1;
    int8_t pl;
};

__device__ __forceinline__ bool master_setup(int i, const int32_t* __restrict__ node_seg,
                                             const int32_t* __restrict__ seg_node_ptr,
                                             const int32_t* __restrict__ master_end, const double* __restrict__ stats,
                                             const double* __restrict__ uk, const int8_t* __restrict__ pl,
                                             const int32_t* __restrict__ node_slot, const int32_t* __restrict__ adj_ptr,
                                             Master& m)
{
    if (node_slot[i] < 0) return false;
    m.s = node_seg[i];
    if (i >= master_end[m.s]) return false;
    const double* st = stats + (size_t)m.s * PK_NSTAT;
    m.x = uk[3 * i]; m.y = uk[3 * i + 1]; m.z = uk[3 * i + 2];
    const int gx = bin_index((m.x - st[PK_ST_BB]) / st[PK_ST_BB + 2], 10);
    const int gy = bin_index((m.y - st[PK_ST_BB + 1]) / st[PK_ST_BB + 3], 5);
    const double thr = st[PK_ST_GRID + 5 * gx + gy];
    m.t2 = thr * thr;
    if (!(m.t2 > 0.0)) return false;
    const double hc = st[PK_ST_CELL];
    m.cx = (int)((m.x - st[PK_ST_LO]) / hc);
    m.cy = (int)((m.y - st[PK_ST_LO + 1]) / hc);
    m.cz = (int)((m.z - st[PK_ST_LO + 2]) / hc);
    m.nend = seg_node_ptr[m.s + 1];
    m.a0 = adj_ptr[i];
    m.a1 = adj_ptr[i + 1];
    m.pl = pl[i];
    return true;
}

// hash slot of neighbour cell k (0..26) of the master's cell
__device__ __forceinline__ uint32_t neighbour_slot(const Master& m, int k, uint32_t hmask)
{
    return cell_hash(m.s, m.cx + k % 3 - 1, m.cy + (k / 3) % 3 - 1, m.cz + k / 9 - 1, hmask);
}

__device__ __forceinline__ bool accept(const Master& m, int i, int j, const double* __restrict__ uk,
                                       const int8_t* __restrict__ pl, const int32_t* __restrict__ adj_col)
{
    if (j <= i || j >= m.nend || pl[j] != m.pl) return false;
    const double dx = m.x - uk[3 * j], dy = m.y - uk[3 * j + 1], dz = m.z - uk[3 * j + 2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 < m.t2)) return false;
    int lo = m.a0, hi = m.a1;  // element neighbours of i, sorted
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (adj_col[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return !(lo < m.a1 && adj_col[lo] == j);
}

// Count pass: 32 lanes per master, lane k searches neighbour cell k (cells whose hash slot an
// earlier lane already covers are skipped); the lanes' counts are summed by shuffles.
__global__ __launch_bounds__(256) void k_pk_count(int N, const int32_t* __restrict__ node_seg,
                                                  const int32_t* __restrict__ seg_node_ptr,
                                                  const int32_t* __restrict__ master_end,
                                                  const double* __restrict__ stats, const double* __restrict__ uk,
                                                  const int8_t* __restrict__ pl, const int32_t* __restrict__ node_slot,
                                                  const int32_t* __restrict__ hstart, const int32_t* __restrict__ cell_nodes,
                                                  const int32_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj_col,
                                                  uint32_t hmask, int32_t* __restrict__ pcount)
{
    const int i = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5);
    const int lane = threadIdx.x & 31;
    if (i >= N) return;  // uniform over the 32 lanes of a master
    Master m;
    if (!master_setup(i, node_seg, seg_node_ptr, master_end, stats, uk, pl, node_slot, adj_ptr, m)) {
        if (lane == 0) pcount[i] = 0;
        return;
    }
    const uint32_t sl = neighbour_slot(m, min(lane, 26), hmask);
    bool dup = lane >= 27;
    for (int k = 0; k < 27; ++k) {
        const uint32_t o = __shfl(sl, k, 32);
        dup |= (k < lane) && (o == sl);
    }
    int cnt = 0;
    if (!dup) {
        const int q1 = hstart[sl + 1];
        for (int q = hstart[sl]; q < q1; ++q) cnt += accept(m, i, cell_nodes[q], uk, pl, adj_col) ? 1 : 0;
    }
    for (int off = 16; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 32);
    if (lane == 0) pcount[i] = cnt;
}

// Fill pass (only when pairs exist): one lane per master collects its slaves, sorts them and
// writes its pairs at its offset, so that the list is ordered by master, then slave.
__global__ __launch_bounds__(256) void k_pk_fill(int N, const int32_t* __restrict__ node_seg,
                                                 const int32_t* __restrict__ seg_node_ptr,
                                                 const int32_t* __restrict__ master_end,
                                                 const double* __restrict__ stats, const double* __restrict__ uk,
                                                 const int8_t* __restrict__ pl, const int32_t* __restrict__ node_slot,
                                                 const int32_t* __restrict__ hstart, const int32_t* __restrict__ cell_nodes,
                                                 const int32_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj_col,
                                                 uint32_t hmask, const int32_t* __restrict__ pstart, int32_t* __restrict__ pa,
                                                 int32_t* __restrict__ pb, double* __restrict__ pw,
                                                 int32_t* __restrict__ status)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int base = pstart[i], room = pstart[i + 1] - base;
    if (room == 0) return;
    Master m;
    if (!master_setup(i, node_seg, seg_node_ptr, master_end, stats, uk, pl, node_slot, adj_ptr, m)) return;
    int js[PK_MAX_PARTNERS];
    int cnt = 0;
    uint32_t slots[27];
    for (int k = 0; k < 27; ++k) {
        const uint32_t sl = neighbour_slot(m, k, hmask);
        bool dup = false;
        for (int q = 0; q < k; ++q) dup |= slots[q] == sl;
        slots[k] = sl;
        if (dup) continue;
        const int q1 = hstart[sl + 1];
        for (int q = hstart[sl]; q < q1; ++q) {
            const int j = cell_nodes[q];
            if (!accept(m, i, j, uk, pl, adj_col)) continue;
            if (cnt < PK_MAX_PARTNERS) js[cnt] = j;
            cnt++;
        }
    }
    if (cnt > PK_MAX_PARTNERS) {
        atomicMax(status + 2, cnt);
        cnt = PK_MAX_PARTNERS;
    }
    cnt = min(cnt, room);  // never past the room the count pass reserved
    for (int k = 1; k < cnt; ++k) {  // ascending slave order
        const int v = js[k];
        int q = k - 1;
        while (q >= 0 && js[q] > v) { js[q + 1] = js[q]; --q; }
        js[q + 1] = v;
    }
    for (int k = 0; k < cnt; ++k) {
        const int j = js[k];
        const double dx = m.x - uk[3 * j], dy = m.y - uk[3 * j + 1], dz = m.z - uk[3 * j + 2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        pa[base + k] = i;
        pb[base + k] = j;
        pw[3 * (size_t)(base + k)] = dx / d;
        pw[3 * (size_t)(base + k) + 1] = dy / d;
        pw[3 * (size_t)(base + k) + 2] = dz / d;
    }
}

// ---- exclusive scan of int32 (n <= SCAN_TILE^2): out[0..n], out[n] = total ---------------------
__global__ __launch_bounds__(SCAN_WG) void k_scan_tiles(const int32_t* __restrict__ in, int n, int32_t* __restrict__ out,
                                                        int32_t* __restrict__ tile_sum)
{
    __shared__ int32_t sh[SCAN_WG];
    const int t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)t * SCAN_PER;
    int32_t v[SCAN_PER], sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        v[k] = base + k < (size_t)n ? in[base + k] : 0;
        sum += v[k];
    }
    sh[t] = sum;
    __syncthreads();
    for (int off = 1; off < SCAN_WG; off <<= 1) {
        const int32_t x = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    int32_t run = sh[t] - sum;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        if (base + k < (size_t)n) out[base + k] = run;
        run += v[k];
    }
    if (t == SCAN_WG - 1) tile_sum[blockIdx.x] = sh[t];
}

__global__ __launch_bounds__(SCAN_WG) void k_scan_top(int32_t* __restrict__ tile_sum, int n_tiles)
{
    __shared__ int32_t sh[SCAN_WG];
    const int t = threadIdx.x;
    int32_t v[SCAN_PER], sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        const int i = t * SCAN_PER + k;
        v[k] = i < n_tiles ? tile_sum[i] : 0;
        sum += v[k];
    }
    sh[t] = sum;
    __syncthreads();
    for (int off = 1; off < SCAN_WG; off <<= 1) {
        const int32_t x = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    int32_t run = sh[t] - sum;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        const int i = t * SCAN_PER + k;
        if (i <= n_tiles) tile_sum[i] = run;  // tile_sum[n_tiles] = total
        run += v[k];
    }
}

__global__ __launch_bounds__(SCAN_WG) void k_scan_add(int32_t* __restrict__ out, int n, const int32_t* __restrict__ tile_sum,
                                                      int n_tiles)
{
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_PER;
    const int32_t add = tile_sum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
        if (base + k < (size_t)n) out[base + k] += add;
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_sum[n_tiles];
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_copy_i32(int32_t* dst, const int32_t* src, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---- forces and Jacobian of one pair (SC/Simulation.cpp:7149-7203, :7205-7295) -------------
// Component k: d = x_a,k - x_b,k, negated when w_k > 0; s = 1 / (1 + exp(-(S/T) d));
// F_k = M mbar s w_k (on a; -F_k on b); D_k = dF_k/dx_a,k = s (1 - s) M mbar w_k (S/T), negated
// when w_k > 0.  The reference's K gets K[a][a] -= D, K[a][b] += D, K[b][b] -= D, K[b][a] += D.
__device__ void pair_eval(int p, const int32_t* __restrict__ pa, const int32_t* __restrict__ pb,
                          const double* __restrict__ pw, const int32_t* __restrict__ node_seg,
                          const double* __restrict__ stats, const double* __restrict__ mass,
                          const double* __restrict__ uk, double F[3], double D[3])
{
    const int a = pa[p], b = pb[p];
    const double c = PK_SAT / stats[(size_t)node_seg[a] * PK_NSTAT + PK_ST_T];
    const double mbar = 0.5 * (mass[a] + mass[b]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double w = pw[3 * (size_t)p + k];
        double d = uk[3 * a + k] - uk[3 * b + k];
        if (w > 0) d *= -1.0;
        const double ex = exp(c * (-1.0 * d));
        F[k] = PK_MULT * mbar / (1 + ex) * w;
        const double sg = 1 / (1 + ex);
        double dd = sg * (1 - sg) * PK_MULT * mbar * w * c;
        if (w > 0) dd *= -1.0;
        D[k] = dd;
    }
}

__global__ __launch_bounds__(256) void k_pk_nodes(int nn, const int32_t* __restrict__ nodes, const int32_t* __restrict__ nptr,
                                                  const int32_t* __restrict__ nref, const int32_t* __restrict__ pa,
                                                  const int32_t* __restrict__ pb, const double* __restrict__ pw,
                                                  const int32_t* __restrict__ node_seg, const double* __restrict__ stats,
                                                  const double* __restrict__ mass, const double* __restrict__ uk,
                                                  double* __restrict__ pf, double* __restrict__ pd)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    double f[3] = {0, 0, 0}, dg[3] = {0, 0, 0};
    for (int r = nptr[k]; r < nptr[k + 1]; ++r) {
        const int ref = nref[r];
        const int p = ref >= 0 ? ref : ~ref;
        double F[3], D[3];
        pair_eval(p, pa, pb, pw, node_seg, stats, mass, uk, F, D);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            f[c] += ref >= 0 ? F[c] : -F[c];
            dg[c] -= D[c];
        }
    }
    const int n = nodes[k];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        pf[3 * n + c] = f[c];
        pd[3 * n + c] = dg[c];
    }
}

// Off-diagonal entries after k_gather; entries in a fixed row or column stay as the fixed-DoF
// rule left them (SC/NewtonRaphsonSolver.cpp:237-255 runs after the packing Jacobian).
__global__ __launch_bounds__(256) void k_pk_offdiag(int np, const int32_t* __restrict__ pa, const int32_t* __restrict__ pb,
                                                    const double* __restrict__ pw, const int32_t* __restrict__ slot,
                                                    const int32_t* __restrict__ node_seg, const double* __restrict__ stats,
                                                    const double* __restrict__ mass, const double* __restrict__ uk,
                                                    const uint8_t* __restrict__ fixed, int nb, double* __restrict__ val,
                                                    float* __restrict__ val32)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double F[3], D[3];
    pair_eval(p, pa, pb, pw, node_seg, stats, mass, uk, F, D);
    const int a = pa[p], b = pb[p];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (fixed[3 * a + k] || fixed[3 * b + k]) continue;
        for (int e = 0; e < 2; ++e) {
            const size_t idx = (size_t)(4 * k) * nb + slot[2 * p + e];
            const double v = val[idx] + D[k];
            val[idx] = v;
            if (val32) val32[idx] = (float)v;
        }
    }
}

void scan_i32(tf_handle* h, const int32_t* in, int n, int32_t* out)
{
    const int tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles >= SCAN_TILE) throw std::runtime_error("packing scan too long");
    PackingState& pk = h->pk;
    if (tiles == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, h->stream, out, 0, 1);
        TF_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_WG), 0, h->stream, in, n, out, pk.scan_tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_WG), 0, h->stream, pk.scan_tmp, tiles);
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(SCAN_WG), 0, h->stream, out, n, pk.scan_tmp, tiles);
    TF_HIP(hipGetLastError());
}

// Prism::calculateReferenceVolume (SC/Prism.cpp:876-908) with
// ShapeBase::calculateVolumeForInputShapeStructure (SC/ShapeBase.cpp:3362-3397): eight triangles
// (the caps and two per lateral quad) each close a tetrahedron with the centroid.
double reference_volume(const double* X)
{
    static const int tri[8][3] = {{0, 1, 2}, {3, 4, 5}, {0, 1, 3}, {1, 3, 4}, {1, 2, 4}, {2, 4, 5}, {0, 2, 3}, {2, 3, 5}};
    double c[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 6; ++a)
        for (int j = 0; j < 3; ++j) c[j] += X[3 * a + j];
    for (int j = 0; j < 3; ++j) c[j] /= 6;
    double V = 0.0;
    for (int t = 0; t < 8; ++t) {
        const double *p0 = X + 3 * tri[t][0], *p1 = X + 3 * tri[t][1], *p2 = X + 3 * tri[t][2];
        double v1[3], v2[3], vm[3];
        for (int j = 0; j < 3; ++j) {
            v1[j] = p1[j] - p0[j];
            v2[j] = p2[j] - p0[j];
            vm[j] = c[j] - p0[j];
        }
        const double b[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
        const double nb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        const double height = std::fabs((vm[0] * b[0] + vm[1] * b[1] + vm[2] * b[2]) / nb);
        V += 1.0 / 3.0 * (height * (nb / 2));
    }
    return V;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
void packing_init(tf_handle* h, const int8_t* node_placement, const uint8_t* exposed, const double* ref_pos_aos)
{
    const int N = h->N, E = h->E;
    std::vector<int8_t> pl(N, 2);
    if (node_placement) {
        for (int n = 0; n < N; ++n) {
            if (node_placement[n] < 0 || node_placement[n] > 3)
                throw std::runtime_error("node placement must be 0 (basal), 1 (apical), 2 (mid) or 3 (lateral)");
            pl[n] = node_placement[n];
        }
    } else {
        // from the exposed faces: corners of an exposed basal cap are basal, of an apical cap apical
        for (int e = 0; e < E; ++e)
            for (int a = 0; a < 6; ++a) {
                const int bit = a < 3 ? 1 : 2;
                if (exposed && (exposed[e] & bit)) pl[h->h_conn[6 * e + a]] = a < 3 ? 0 : 1;
            }
    }
    // element placement (ShapeBase::setTissuePlacement, SC/ShapeBase.cpp:718-779): basal side
    // lengths count for basal elements and elements spanning the tissue
    std::vector<uint8_t> basal(E, 0);
    std::vector<double> vref(E);
    for (int e = 0; e < E; ++e) {
        bool ba = false, lat = false;
        for (int a = 0; a < 6; ++a) {
            const int p = pl[h->h_conn[6 * e + a]];
            ba |= p == 0;
            lat |= p == 3;
        }
        basal[e] = !lat && ba;
        vref[e] = reference_volume(ref_pos_aos + (size_t)18 * e);
    }
    PackingState& pk = h->pk;
    pk.h_node_pl = pl;
    pk.node_pl = pk_alloc<int8_t>(N);
    pk.elem_basal = pk_alloc<uint8_t>(E);
    pk.vref = pk_alloc<double>(E);
    pk.mass = pk_alloc<double>(N);
    pk.vpn = pk_alloc<double>(E);
    pk.seg_stats = pk_alloc<double>((size_t)h->S * PK_NSTAT);
    pk.f = pk_alloc<double>((size_t)3 * N);
    pk.d = pk_alloc<double>((size_t)3 * N);
    pk_h2d(h, pk.node_pl, pl.data(), N);
    pk_h2d(h, pk.elem_basal, basal.data(), E);
    pk_h2d(h, pk.vref, vref.data(), E);
    std::vector<double> z((size_t)std::max<int64_t>((int64_t)3 * N, (int64_t)h->S * PK_NSTAT), 0.0);
    pk_h2d(h, pk.mass, z.data(), N);
    pk_h2d(h, pk.seg_stats, z.data(), (size_t)h->S * PK_NSTAT);
    pk_h2d(h, pk.f, z.data(), (size_t)3 * N);
    pk_h2d(h, pk.d, z.data(), (size_t)3 * N);
}

void packing_free(tf_handle* h)
{
    PackingState& pk = h->pk;
    pk_free(pk.node_pl); pk_free(pk.elem_basal); pk_free(pk.vref); pk_free(pk.mass); pk_free(pk.seg_stats);
    pk_free(pk.f); pk_free(pk.d); pk_free(pk.adj_ptr); pk_free(pk.adj_col); pk_free(pk.node_seg);
    pk_free(pk.master_end); pk_free(pk.chunks); pk_free(pk.seg_chunk_ptr); pk_free(pk.chunk_part);
    pk_free(pk.runs); pk_free(pk.seg_run_ptr); pk_free(pk.run_part); pk_free(pk.ev); pk_free(pk.vpn); pk_free(pk.hcnt);
    pk_free(pk.hstart); pk_free(pk.cell_nodes); pk_free(pk.node_slot); pk_free(pk.pcount); pk_free(pk.pstart);
    pk_free(pk.scan_tmp); pk_free(pk.a); pk_free(pk.b); pk_free(pk.w); pk_free(pk.slot); pk_free(pk.nodes);
    pk_free(pk.nptr); pk_free(pk.nref);
    if (pk.ev_count) { (void)hipEventDestroy(pk.ev_count); pk.ev_count = nullptr; }
    if (pk.graph) { (void)hipGraphExecDestroy(pk.graph); pk.graph = nullptr; }
}

// Detection work arrays and node tables, allocated on first use.
void packing_prepare(tf_handle* h)
{
    PackingState& pk = h->pk;
    if (pk.adj_ptr) return;
    const int N = h->N, E = h->E, S = h->S;
    // element neighbours of every node, sorted (the element pattern without bound or pair blocks)
    std::vector<int32_t> aptr(N + 1, 0), acol;
    acol.reserve((size_t)N * 20);
    std::vector<int32_t> row;
    for (int A = 0; A < N; ++A) {
        row.clear();
        for (int k = h->h_node_ptr[A]; k < h->h_node_ptr[A + 1]; ++k) {
            const int e = h->h_node_src[k] / 6;
            for (int b = 0; b < 6; ++b) row.push_back(h->h_conn[6 * e + b]);
        }
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        acol.insert(acol.end(), row.begin(), row.end());
        aptr[A + 1] = (int32_t)acol.size();
    }
    std::vector<int32_t> nseg(N), mend(S);
    for (int s = 0; s < S; ++s) {
        const int n0 = h->h_node_seg_ptr[s], n1 = h->h_node_seg_ptr[s + 1];
        for (int n = n0; n < n1; ++n) nseg[n] = s;
        // parallellisation chunks of detectPacingNodes (SC/Simulation.cpp:8204-8218, :8422-8437)
        const int nn = n1 - n0;
        const int seg_size = std::max(1, nn / 16);
        const int n_chunks = nn / seg_size;
        mend[s] = n0 + (n_chunks - 1) * seg_size;
    }
    int H = 1024;
    while (H < 2 * N) H <<= 1;
    pk.H = H;
    pk.adj_ptr = pk_alloc<int32_t>(N + 1);
    pk.adj_col = pk_alloc<int32_t>(acol.size());
    pk.node_seg = pk_alloc<int32_t>(N);
    pk.master_end = pk_alloc<int32_t>(S);
    // runs of PK_CHUNK elements inside every tissue (the grid-bin partial sums)
    std::vector<int4> chunks;
    std::vector<int32_t> scp(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        for (int e = h->h_elem_seg_ptr[s]; e < h->h_elem_seg_ptr[s + 1]; e += PK_CHUNK)
            chunks.push_back(int4{s, e, std::min(e + PK_CHUNK, h->h_elem_seg_ptr[s + 1]), 0});
        scp[s + 1] = (int32_t)chunks.size();
    }
    pk.n_chunks = (int32_t)chunks.size();
    // runs of PK_NODE_RUN nodes inside every tissue (the bounding-box partials); an empty tissue
    // gets one empty run
    std::vector<int4> runs;
    std::vector<int32_t> srp(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        const int n0 = h->h_node_seg_ptr[s], n1 = h->h_node_seg_ptr[s + 1];
        if (n1 == n0) runs.push_back(int4{s, n0, n0, 0});
        for (int n = n0; n < n1; n += PK_NODE_RUN) runs.push_back(int4{s, n, std::min(n + PK_NODE_RUN, n1), 0});
        srp[s + 1] = (int32_t)runs.size();
    }
    pk.n_runs = (int32_t)runs.size();
    pk.runs = pk_alloc<int4>(runs.size());
    pk.seg_run_ptr = pk_alloc<int32_t>(S + 1);
    pk.run_part = pk_alloc<double>(runs.size() * 10);
    pk.chunks = pk_alloc<int4>(chunks.size());
    pk.seg_chunk_ptr = pk_alloc<int32_t>(S + 1);
    pk.chunk_part = pk_alloc<double>(chunks.size() * PK_NPART);
    pk.ev = pk_alloc<double>((size_t)5 * E);
    pk.hcnt = pk_alloc<int32_t>(H);
    pk.hstart = pk_alloc<int32_t>(H + 1);
    pk.cell_nodes = pk_alloc<int32_t>(N);
    pk.node_slot = pk_alloc<int32_t>(N);
    pk.pcount = pk_alloc<int32_t>(N);
    pk.pstart = pk_alloc<int32_t>(N + 1);
    pk.scan_tmp = pk_alloc<int32_t>((std::max(H, N) + SCAN_TILE - 1) / SCAN_TILE + 1);
    pk_h2d(h, pk.adj_ptr, aptr.data(), N + 1);
    pk_h2d(h, pk.adj_col, acol.data(), acol.size());
    pk_h2d(h, pk.node_seg, nseg.data(), N);
    pk_h2d(h, pk.master_end, mend.data(), S);
    pk_h2d(h, pk.chunks, chunks.data(), chunks.size());
    pk_h2d(h, pk.seg_chunk_ptr, scp.data(), S + 1);
    pk_h2d(h, pk.runs, runs.data(), runs.size());
    pk_h2d(h, pk.seg_run_ptr, srp.data(), S + 1);
}

void launch_packing_mass(tf_handle* h)
{
    hipLaunchKernelGGL(k_pk_vpn, dim3((h->E + 255) / 256), dim3(256), 0, h->stream, h->fg, h->pk.vref, h->E, h->pk.vpn);
    hipLaunchKernelGGL(k_pk_mass, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->node_ptr, h->node_src,
                       h->pk.vpn, h->ablated, h->N, h->pk.mass);
    TF_HIP(hipGetLastError());
}

void packing_reserve(tf_handle* h, int n)
{
    PackingState& pk = h->pk;
    if (n <= pk.cap && pk.a) return;
    const int cap = std::max({n, 2 * pk.cap, 256});
    pk_free(pk.a); pk_free(pk.b); pk_free(pk.w); pk_free(pk.slot);
    pk.a = pk_alloc<int32_t>(cap);
    pk.b = pk_alloc<int32_t>(cap);
    pk.w = pk_alloc<double>((size_t)3 * cap);
    pk.slot = pk_alloc<int32_t>((size_t)2 * cap);
    pk.cap = cap;
}

// Pattern, slots and node lists for the pairs now in pk.a / pk.b (host copies a, b).
void packing_install(tf_handle* h, const std::vector<int32_t>& a, const std::vector<int32_t>& b)
{
    PackingState& pk = h->pk;
    const int P = (int)a.size(), N = h->N;
    if (P == 0 && pk.n == 0) return;  // nothing in force before or after
    auto find_blk = [&](int A, int B) -> int {
        const int32_t* c0 = h->h_col.data() + h->h_row_ptr[A];
        const int32_t* c1 = h->h_col.data() + h->h_row_ptr[A + 1];
        const int32_t* it = std::lower_bound(c0, c1, B);
        return (it != c1 && *it == B) ? (int)(it - h->h_col.data()) : -1;
    };
    bool missing = false;
    for (int p = 0; p < P && !missing; ++p) missing = find_blk(a[p], b[p]) < 0;
    if (missing) {
        // the pattern keeps the blocks of earlier pairs while they are few, so that contacts
        // that open and close do not rebuild it every step
        std::vector<int64_t> keys;
        for (int p = 0; p < P; ++p)
            keys.push_back((int64_t)std::min(a[p], b[p]) * N + std::max(a[p], b[p]));
        if (pk.h_blocks.size() <= 2 * (size_t)P + 64) keys.insert(keys.end(), pk.h_blocks.begin(), pk.h_blocks.end());
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        pk.h_blocks = keys;
        build_pattern(h);
        build_tiles(h, h->wg, h->tile_cap);
    }
    std::vector<int32_t> slot((size_t)2 * P);
    for (int p = 0; p < P; ++p) {
        slot[2 * p] = find_blk(a[p], b[p]);
        slot[2 * p + 1] = find_blk(b[p], a[p]);
        if (slot[2 * p] < 0 || slot[2 * p + 1] < 0) throw std::runtime_error("packing pair missing from the pattern");
    }
    // nodes with pairs and their pair references in pair order (p: node a of pair p, ~p: node b)
    std::vector<int32_t> cnt(N, 0);
    for (int p = 0; p < P; ++p) { cnt[a[p]]++; cnt[b[p]]++; }
    std::vector<int32_t> nodes, nptr{0}, pos(N, -1);
    for (int n = 0; n < N; ++n)
        if (cnt[n]) {
            pos[n] = nptr.back();
            nodes.push_back(n);
            nptr.push_back(nptr.back() + cnt[n]);
        }
    std::vector<int32_t> nref((size_t)2 * P);
    for (int p = 0; p < P; ++p) {
        nref[pos[a[p]]++] = p;
        nref[pos[b[p]]++] = ~p;
    }
    if (P) {
        if ((int)nodes.size() > pk.node_cap || 2 * P > pk.ref_cap) {  // node lists grow only
            pk_free(pk.nodes); pk_free(pk.nptr); pk_free(pk.nref);
            pk.node_cap = std::max<int>(2 * pk.node_cap, (int)nodes.size());
            pk.ref_cap = std::max<int>(2 * pk.ref_cap, 2 * P);
            pk.nodes = pk_alloc<int32_t>(pk.node_cap);
            pk.nptr = pk_alloc<int32_t>(pk.node_cap + 1);
            pk.nref = pk_alloc<int32_t>(pk.ref_cap);
        }
        pk_h2d(h, pk.slot, slot.data(), slot.size());
        pk_h2d(h, pk.nodes, nodes.data(), nodes.size());
        pk_h2d(h, pk.nptr, nptr.data(), nptr.size());
        pk_h2d(h, pk.nref, nref.data(), nref.size());
        launch_fill(h, pk.f, 0.0, 3 * (int64_t)N);
        launch_fill(h, pk.d, 0.0, 3 * (int64_t)N);
    }
    pk.n = P;
    pk.nn = (int32_t)nodes.size();
    pk.h_a = a;
    pk.h_b = b;
}

// the reference's bounding box of every tissue at the current positions into stats (S x PK_NSTAT)
void launch_reference_box(tf_handle* h, double* stats)
{
    packing_prepare(h);
    PackingState& pk = h->pk;
    hipLaunchKernelGGL(k_pk_box, dim3(pk.n_runs), dim3(256), 0, h->stream, pk.runs, h->seg_node_ptr, h->uk, pk.run_part);
    hipLaunchKernelGGL(k_pk_box_finish, dim3(h->S), dim3(256), 0, h->stream, pk.seg_run_ptr, h->seg_node_ptr, pk.runs,
                       pk.run_part, h->uk, stats);
    TF_HIP(hipGetLastError());
}

// The detection's kernels up to the pair count (fixed arguments for the handle's lifetime: device
// tables allocated once, sizes of the mesh)
static void packing_detect_kernels(tf_handle* h)
{
    PackingState& pk = h->pk;
    const int N = h->N, S = h->S;
    const uint32_t hmask = (uint32_t)pk.H - 1u;
    launch_packing_mass(h);
    hipLaunchKernelGGL(k_pk_evals, dim3((h->E + 255) / 256), dim3(256), 0, h->stream, h->uk, h->conn, h->E, pk.elem_basal,
                       pk.ev);
    launch_reference_box(h, pk.seg_stats);
    if (pk.n_chunks > 0)
        hipLaunchKernelGGL(k_pk_bins, dim3(pk.n_chunks), dim3(PK_CHUNK), 0, h->stream, pk.chunks, pk.ev, h->E,
                           pk.seg_stats, pk.chunk_part);
    hipLaunchKernelGGL(k_pk_finish, dim3(S), dim3(1024), 0, h->stream, pk.seg_chunk_ptr, pk.chunk_part, pk.seg_stats);
    hipLaunchKernelGGL(k_fill_i32, dim3((pk.H + 255) / 256), dim3(256), 0, h->stream, pk.hcnt, 0, pk.H);
    hipLaunchKernelGGL(k_pk_cells, dim3((N + 255) / 256), dim3(256), 0, h->stream, N, pk.node_seg, pk.seg_stats, h->uk,
                       pk.mass, pk.node_pl, hmask, pk.node_slot, pk.hcnt);
    TF_HIP(hipGetLastError());
    scan_i32(h, pk.hcnt, pk.H, pk.hstart);
    hipLaunchKernelGGL(k_copy_i32, dim3((pk.H + 255) / 256), dim3(256), 0, h->stream, pk.hcnt, pk.hstart, pk.H);
    hipLaunchKernelGGL(k_pk_fill_cells, dim3((N + 255) / 256), dim3(256), 0, h->stream, N, pk.node_slot, pk.hcnt,
                       pk.cell_nodes);
    hipLaunchKernelGGL(k_pk_count, dim3((unsigned)(((size_t)N * 32 + 255) / 256)), dim3(256), 0, h->stream, N, pk.node_seg,
                       h->seg_node_ptr, pk.master_end, pk.seg_stats, h->uk, pk.node_pl, pk.node_slot, pk.hstart,
                       pk.cell_nodes, pk.adj_ptr, pk.adj_col, hmask, pk.pcount);
    TF_HIP(hipGetLastError());
    scan_i32(h, pk.pcount, N, pk.pstart);
}

void packing_detect_launch(tf_handle* h)
{
    packing_prepare(h);
    PackingState& pk = h->pk;
    const int N = h->N;
    if (h->opts.use_graph) {  // the 17 small kernels from one graph: one launch for the host
        if (!pk.graph) {
            hipGraph_t gr;
            TF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            try {
                packing_detect_kernels(h);
            } catch (...) {
                (void)hipStreamEndCapture(h->stream, &gr);
                throw;
            }
            TF_HIP(hipStreamEndCapture(h->stream, &gr));
            const hipError_t e = hipGraphInstantiate(&pk.graph, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) pk.graph = nullptr;
            TF_HIP(e);
        }
        TF_HIP(hipGraphLaunch(pk.graph, h->stream));
    } else {
        packing_detect_kernels(h);
    }
    if (!pk.ev_count) TF_HIP(hipEventCreateWithFlags(&pk.ev_count, hipEventDisableTiming));
    TF_HIP(hipMemcpyAsync(h->h_status + 8, pk.pstart + N, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    TF_HIP(hipEventRecord(pk.ev_count, h->stream));
}

int packing_detect_finish(tf_handle* h)
{
    PackingState& pk = h->pk;
    const int N = h->N;
    const uint32_t hmask = (uint32_t)pk.H - 1u;
    TF_HIP(hipEventSynchronize(pk.ev_count));
    const int32_t total = h->h_status[8];
    std::vector<int32_t> a(total), b(total);
    if (total > 0) {
        packing_reserve(h, total);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, h->stream, h->status + 2, 0, 1);
        hipLaunchKernelGGL(k_pk_fill, dim3((N + 255) / 256), dim3(256), 0, h->stream, N, pk.node_seg, h->seg_node_ptr,
                           pk.master_end, pk.seg_stats, h->uk, pk.node_pl, pk.node_slot, pk.hstart, pk.cell_nodes,
                           pk.adj_ptr, pk.adj_col, hmask, pk.pstart, pk.a, pk.b, pk.w, h->status);
        TF_HIP(hipGetLastError());
        int32_t over = 0;
        TF_HIP(hipMemcpyAsync(&over, h->status + 2, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        TF_HIP(hipMemcpyAsync(a.data(), pk.a, sizeof(int32_t) * total, hipMemcpyDeviceToHost, h->stream));
        TF_HIP(hipMemcpyAsync(b.data(), pk.b, sizeof(int32_t) * total, hipMemcpyDeviceToHost, h->stream));
        TF_HIP(hipStreamSynchronize(h->stream));
        if (over > 0)
            throw std::runtime_error("a node has " + std::to_string(over) + " packing partners (at most " +
                                     std::to_string(PK_MAX_PARTNERS) + " are supported)");
    }
    packing_install(h, a, b);
    return total;
}

int packing_detect(tf_handle* h)
{
    packing_detect_launch(h);
    return packing_detect_finish(h);
}

void packing_clear(tf_handle* h)
{
    if (h->pk.n == 0) return;
    packing_install(h, {}, {});
}

void launch_packing_nodes(tf_handle* h)
{
    PackingState& pk = h->pk;
    if (!pk.n) return;
    hipLaunchKernelGGL(k_pk_nodes, dim3((pk.nn + 255) / 256), dim3(256), 0, h->stream, pk.nn, pk.nodes, pk.nptr, pk.nref,
                       pk.a, pk.b, pk.w, pk.node_seg, pk.seg_stats, pk.mass, h->uk, pk.f, pk.d);
    TF_HIP(hipGetLastError());
}

void launch_packing_offdiag(tf_handle* h)
{
    PackingState& pk = h->pk;
    if (!pk.n) return;
    hipLaunchKernelGGL(k_pk_offdiag, dim3((pk.n + 255) / 256), dim3(256), 0, h->stream, pk.n, pk.a, pk.b, pk.w, pk.slot,
                       pk.node_seg, pk.seg_stats, pk.mass, h->uk, h->fixed, h->nb, h->val,
                       fp32_values(h) ? h->val32 : nullptr);
    TF_HIP(hipGetLastError());
}
