// tf_growth.hip -- growth functions evaluated on the device every growth update.
//
// Simulation::calculateGrowth (SC/Simulation.cpp:9651-9670) starts every element's growth
// increment at the identity and left-multiplies the increment of every growth function active at
// the step's start time (ShapeBase::updateGrowthIncrement, SC/ShapeBase.cpp:567-586):
//   kind 1, uniform (calculateGrowthUniform, SC/Simulation.cpp:9757-9799): R diag(exp(r dt)) R^T,
//           R the rotation by the function's xy angle (calculateFgFromRates, SC/ShapeBase.cpp:396-421);
//   kind 2, ring (calculateGrowthRing, SC/Simulation.cpp:9801-9898): the same with the rate scaled
//           by 1 - (d - r_in) / (r_out - r_in) for elements whose centre lies strictly between the
//           radii, in the float arithmetic the reference uses for d and the scale;
//   kind 3, grid (calculateGrowthGridBased, SC/Simulation.cpp:10065-10117): rates and xy angles on
//           an nx x ny grid over the tissue's bounding box, at the element's relative position
//           (pinned on an earlier mesh or current), from the grid cell's four corners, nearest
//           corner or bilinear (calculateFgFromGridCorners, SC/ShapeBase.cpp:432-565), with the
//           corners' angles first brought next to each other (the grid's
//           preCalculateAnglesForCompatibleAveraging, GrowthFunctionTypes.h:434-587) and the
//           z-range test on the element's initial relative height (isGrowthRateApplicable,
//           SC/ShapeBase.cpp:277-300).
// The product goes to the increment array k_grow applies (with the growth-rotation rule).
// Relative positions: (centre - box min) / box size in the reference's bounding box
// (Simulation::calculateBoundingBox, SC/Simulation.cpp:9148-9189), the element's column's apical
// value copied down to the first layers - 2 elements of the column (rel_src; the reference's
// updateRelativePositionsToApicalPositioning).
#include "tf_internal.hpp"
#include "tf_element.hpp"

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int GF_CELL = 20;  // per grid cell: growth of the 4 corners (12), angles (4), eliminated (4)

struct DevGrowthFn {
    int32_t kind, nx, ny, cell0;  // cell0: the grid's first cell in the cell table
    double rate[3];
    double c, s;                  // cos and sin of the xy angle (uniform, ring)
    double cx, cy;                // ring centre (float values)
    float rin, rout;              // ring radii (float values)
    double zmin, zmax;            // grid: relative-height range
};

struct ActiveFns {
    int32_t n;
    int32_t idx[TF_GROWTH_MAX_ACTIVE];
};

// R diag(exp(r dt)) R^T, R the rotation by the angle with cosine c and sine s about z
__device__ void rate_increment(double r0, double r1, double r2, double dt, double c, double s, double* out)
{
    const double d0 = exp(r0 * dt), d1 = exp(r1 * dt), d2 = exp(r2 * dt);
    out[0] = c * d0 * c + s * d1 * s;
    out[1] = c * d0 * s - s * d1 * c;
    out[2] = 0.0;
    out[3] = s * d0 * c - c * d1 * s;
    out[4] = s * d0 * s + c * d1 * c;
    out[5] = 0.0;
    out[6] = 0.0;
    out[7] = 0.0;
    out[8] = d2;
}

__device__ void element_centre(const int32_t* __restrict__ conn, const double* __restrict__ pos, int E, int e,
                               double& x, double& y)
{
    x = 0.0;
    y = 0.0;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const int n = conn[(size_t)a * E + e];
        x += pos[3 * n];
        y += pos[3 * n + 1];
    }
    x /= 6;
    y /= 6;
}

// relative x, y of element e's centre (through rel_src) in its tissue's reference bounding box
__device__ void relative_position(const int32_t* __restrict__ conn, const double* __restrict__ pos, int E, int e,
                                  const int32_t* __restrict__ rel_src, const int32_t* __restrict__ elem_seg,
                                  const double* __restrict__ box, double& rx, double& ry)
{
    const int src = rel_src[e];
    double x, y;
    element_centre(conn, pos, E, src, x, y);
    const double* b = box + (size_t)elem_seg[src] * PK_NSTAT + PK_ST_BB;  // min x, min y, size x, size y
    rx = (x - b[0]) / b[2];
    ry = (y - b[1]) / b[3];
}

__global__ __launch_bounds__(256) void k_growth_rel(const int32_t* __restrict__ conn, const double* __restrict__ pos,
                                                    int E, const int32_t* __restrict__ rel_src,
                                                    const int32_t* __restrict__ elem_seg, const double* __restrict__ box,
                                                    double* __restrict__ rel)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double rx, ry;
    relative_position(conn, pos, E, e, rel_src, elem_seg, box, rx, ry);
    rel[e] = rx;
    rel[(size_t)E + e] = ry;
}

// ShapeBase::convertRelativePosToGridIndex (SC/ShapeBase.cpp:633-660): cell and fractions
__device__ void grid_index(double rel, int n, int& i, double& f)
{
    const double p = rel * (double)(float)(n - 1);
    i = (int)floor(p);
    f = p - i;
    if (i >= n - 1) { i = n - 2; f = 1.0; }
    else if (i < 0) { i = 0; f = 0.0; }
}

__global__ __launch_bounds__(256) void k_growth_eval(const int32_t* __restrict__ conn, const double* __restrict__ pos,
                                                     int E, const DevGrowthFn* __restrict__ fns,
                                                     const double* __restrict__ cells, ActiveFns act,
                                                     const double* __restrict__ rel_pinned,
                                                     const double* __restrict__ z_init,
                                                     const int32_t* __restrict__ rel_src,
                                                     const int32_t* __restrict__ elem_seg,
                                                     const double* __restrict__ box, int pinned, int interpolation,
                                                     double dt, double* __restrict__ ginc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double tot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int a = 0; a < act.n; ++a) {
        const DevGrowthFn f = fns[act.idx[a]];
        double inc[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (f.kind == 1) {
            rate_increment(f.rate[0], f.rate[1], f.rate[2], dt, f.c, f.s, inc);
        } else if (f.kind == 2) {
            double x, y;
            element_centre(conn, pos, E, e, x, y);
            const double dx = f.cx - x, dy = f.cy - y;
            const double d2 = dx * dx + dy * dy;
            if (d2 > (double)(f.rin * f.rin) && d2 < (double)(f.rout * f.rout)) {
                const float dist = (float)sqrt(d2);
                const double sf = 1.0 - (double)((dist - f.rin) / (f.rout - f.rin));
                rate_increment(f.rate[0] * sf, f.rate[1] * sf, f.rate[2] * sf, dt, f.c, f.s, inc);
            }
        } else if (f.kind == 3 && z_init[e] >= f.zmin && z_init[e] <= f.zmax) {
            double rx, ry;
            if (pinned) {
                rx = rel_pinned[e];
                ry = rel_pinned[(size_t)E + e];
            } else {
                relative_position(conn, pos, E, e, rel_src, elem_seg, box, rx, ry);
            }
            int ix, iy;
            double X, Y;
            grid_index(rx, f.nx, ix, X);
            grid_index(ry, f.ny, iy, Y);
            const double* cl = cells + (size_t)(f.cell0 + ix * (f.ny - 1) + iy) * GF_CELL;
            double g[3], angle = 0.0;
            if (interpolation == 0) {  // the nearest corner, no angle
                const int k = X > 0.5 ? (Y > 0.5 ? 3 : 1) : (Y > 0.5 ? 2 : 0);
                g[0] = cl[3 * k];
                g[1] = cl[3 * k + 1];
                g[2] = cl[3 * k + 2];
            } else {
                const double w[4] = {(1.0 - X) * (1.0 - Y), X * (1.0 - Y), (1.0 - X) * Y, X * Y};
                double frac_elim = 0.0;
                for (int k = 0; k < 4; ++k)
                    if (cl[16 + k] != 0.0) frac_elim += w[k];
                angle = cl[12] * w[0] + cl[13] * w[1] + cl[14] * w[2] + cl[15] * w[3];
                if (frac_elim > 0) angle = frac_elim >= 0.9999999 ? 0.0 : angle / (1.0 - frac_elim);
                for (int d = 0; d < 3; ++d) g[d] = cl[d] * w[0] + cl[3 + d] * w[1] + cl[6 + d] * w[2] + cl[9 + d] * w[3];
            }
            if (angle != 0.0) rate_increment(g[0], g[1], g[2], dt, cos(angle), sin(angle), inc);
            else {
                inc[0] = exp(g[0] * dt);
                inc[4] = exp(g[1] * dt);
                inc[8] = exp(g[2] * dt);
            }
        }
        double t[9];
        tf::mm3(inc, tot, t);
#pragma unroll
        for (int q = 0; q < 9; ++q) tot[q] = t[q];
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) ginc[(size_t)q * E + e] = tot[q];
}

// The grid's preCalculateAnglesForCompatibleAveraging for one cell (corners (i, j), (i+1, j),
// (i, j+1), (i+1, j+1)): corners whose xy growth aspect ratio is below 1.1 are eliminated (angle
// 0); every angle is brought into [0, pi]; from the first kept corner on, each angle is moved by
// a multiple of pi/2 to within pi/4 of that corner's, swapping the x and y growths on odd moves.
void compatible_cell(const double* rates, const double* angles, int ny, int i, int j, double* out)
{
    const double q = M_PI;
    const int ci[4] = {i, i + 1, i, i + 1}, cj[4] = {j, j, j + 1, j + 1};
    double g[4][3], ang[4];
    bool elim[4], flip[4] = {false, false, false, false};
    for (int k = 0; k < 4; ++k) {
        const size_t p = (size_t)ci[k] * ny + cj[k];
        for (int d = 0; d < 3; ++d) g[k][d] = rates[3 * p + d];
        double ar = rates[3 * p] / rates[3 * p + 1];
        if (ar < 0) ar = -ar;
        if (ar < 1.0) ar = 1.0 / ar;
        elim[k] = ar < 1.1;  // NaN keeps the corner, as the reference's comparison does
        ang[k] = elim[k] ? 0.0 : angles[p];
        while (ang[k] < 0) ang[k] += q;
        while (ang[k] > q) ang[k] -= q;
    }
    int sel = 3;
    double tet = 0.0;
    for (int k = 0; k < 4; ++k)
        if (!elim[k]) { tet = ang[k]; sel = k; break; }
    for (int k = sel; k < 4; ++k) {
        const double d = ang[k] - tet;
        if (d > 0) {
            if (q / 4 <= d && d < 3 * q / 4) { ang[k] -= q / 2; flip[k] = true; }
            else if (3 * q / 4 <= d && d < 5 * q / 4) ang[k] -= q;
            else if (5 * q / 4 <= d && d < 7 * q / 4) { ang[k] -= 6 * q / 4; flip[k] = true; }
            else if (d >= 7 * q / 4) ang[k] -= 2 * q;
        } else if (d < 0) {
            if (-3 * q / 4 < d && d <= -q / 4) { ang[k] += q / 2; flip[k] = true; }
            else if (-5 * q / 4 < d && d <= -3 * q / 4) ang[k] += q;
            else if (-7 * q / 4 < d && d <= -5 * q / 4) { ang[k] += 6 * q / 4; flip[k] = true; }
            else if (d <= -7 * q / 4) ang[k] += 2 * q;
        }
    }
    for (int k = 0; k < 4; ++k) {
        out[3 * k] = flip[k] ? g[k][1] : g[k][0];
        out[3 * k + 1] = flip[k] ? g[k][0] : g[k][1];
        out[3 * k + 2] = g[k][2];
        out[12 + k] = ang[k];
        out[16 + k] = elim[k] ? 1.0 : 0.0;
    }
}

template <typename T>
T* gf_upload(tf_handle* h, const std::vector<T>& v)
{
    void* p = nullptr;
    const size_t n = std::max<size_t>(1, v.size());
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) throw tf_hip_error(e, "hipMalloc", __FILE__, __LINE__);
    if (!v.empty()) {
        TF_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
        TF_HIP(hipStreamSynchronize(h->stream));
    }
    return (T*)p;
}

}  // namespace

void growth_functions_free(tf_handle* h)
{
    tf_handle::GrowthFns& gf = h->gf;
    for (void* p : {(void*)gf.fns, (void*)gf.cells, (void*)gf.rel, (void*)gf.z_init, (void*)gf.rel_src, (void*)gf.box})
        if (p) (void)hipFree(p);
    gf = tf_handle::GrowthFns{};
}

// the reference's bounding box of every tissue at the current positions, then (pin) the
// relative positions the pinned grids read
static void growth_box(tf_handle* h)
{
    tf_handle::GrowthFns& gf = h->gf;
    if (!gf.box) gf.box = gf_upload(h, std::vector<double>((size_t)h->S * PK_NSTAT, 0.0));
    launch_reference_box(h, gf.box);
}

void growth_functions_pin(tf_handle* h)
{
    tf_handle::GrowthFns& gf = h->gf;
    if (gf.n_fn == 0) return;
    growth_box(h);
    hipLaunchKernelGGL(k_growth_rel, dim3((h->E + 255) / 256), dim3(256), 0, h->stream, h->conn, h->uk, h->E, gf.rel_src,
                       h->elem_seg, gf.box, gf.rel);
    TF_HIP(hipGetLastError());
}

void growth_functions_set(tf_handle* h, int n, const tf_growth_function* f, const double* z_init,
                          const int32_t* rel_src, int pinned, int interpolation)
{
    const int E = h->E;
    std::vector<DevGrowthFn> fns(n);
    std::vector<double> cells;
    std::vector<double> t0(n), t1(n);
    bool any_grid = false, needs_box = false;
    for (int k = 0; k < n; ++k) {
        const tf_growth_function& g = f[k];
        DevGrowthFn& d = fns[k];
        d = DevGrowthFn{};
        d.kind = g.kind;
        t0[k] = g.t0;
        t1[k] = g.t1;
        for (int i = 0; i < 3; ++i) d.rate[i] = g.rate[i];
        d.c = std::cos(g.angle);
        d.s = std::sin(g.angle);
        if (g.kind == 2) {
            d.cx = (double)(float)g.centre[0];
            d.cy = (double)(float)g.centre[1];
            d.rin = (float)g.inner;
            d.rout = (float)g.outer;
        } else if (g.kind == 3) {
            any_grid = true;
            needs_box = needs_box || !pinned;
            d.nx = g.nx;
            d.ny = g.ny;
            d.zmin = g.zmin;
            d.zmax = g.zmax;
            d.cell0 = (int32_t)(cells.size() / GF_CELL);
            const size_t c0 = cells.size();
            cells.resize(c0 + (size_t)(g.nx - 1) * (g.ny - 1) * GF_CELL);
            for (int i = 0; i < g.nx - 1; ++i)
                for (int j = 0; j < g.ny - 1; ++j)
                    compatible_cell(g.grid_rates, g.grid_angles, g.ny, i, j,
                                    cells.data() + c0 + ((size_t)i * (g.ny - 1) + j) * GF_CELL);
        }
    }
    growth_functions_free(h);
    tf_handle::GrowthFns& gf = h->gf;
    gf.fns = gf_upload(h, fns);
    gf.cells = gf_upload(h, cells);
    gf.t0 = t0;
    gf.t1 = t1;
    gf.kind.resize(n);
    for (int k = 0; k < n; ++k) gf.kind[k] = f[k].kind;
    gf.pinned = pinned;
    gf.interpolation = interpolation;
    gf.needs_box = needs_box;
    gf.z_init = gf_upload(h, std::vector<double>(z_init, z_init + E));
    gf.rel_src = gf_upload(h, std::vector<int32_t>(rel_src, rel_src + E));
    gf.rel = gf_upload(h, std::vector<double>((size_t)2 * E, 0.0));
    gf.n_fn = n;
    if (!h->ginc) h->ginc = gf_upload(h, std::vector<double>((size_t)9 * E, 0.0));
    if (any_grid && pinned) growth_functions_pin(h);  // pinned on the current mesh
}

// The increments of the functions active at the handle's time into ginc (ahead of k_grow).
void growth_functions_eval(tf_handle* h, double dt)
{
    tf_handle::GrowthFns& gf = h->gf;
    ActiveFns act{};
    bool box = false;
    for (int k = 0; k < gf.n_fn; ++k) {
        if (!(gf.t0[k] <= h->time && h->time < gf.t1[k])) continue;
        if (act.n == TF_GROWTH_MAX_ACTIVE)
            throw std::invalid_argument("more than " + std::to_string(TF_GROWTH_MAX_ACTIVE) +
                                        " growth functions active at one time");
        act.idx[act.n++] = k;
        box = box || (gf.kind[k] == 3 && gf.needs_box);
    }
    if (box) growth_box(h);
    hipLaunchKernelGGL(k_growth_eval, dim3((h->E + 255) / 256), dim3(256), 0, h->stream, h->conn, h->uk, h->E,
                       static_cast<const DevGrowthFn*>(gf.fns),
                       gf.cells, act, gf.rel, gf.z_init, gf.rel_src, h->elem_seg, gf.box, gf.pinned, gf.interpolation, dt,
                       h->ginc);
    TF_HIP(hipGetLastError());
}
