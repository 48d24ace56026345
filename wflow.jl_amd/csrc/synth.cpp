// Synthetic catchment generator (host only): depression filling, D8 flow directions and
// upstream-area accumulation on an (x, y) raster with x the fastest index.
//
// This is the workload generator for the synthetic static maps that stand in for the
// netCDF inputs the reference reads (`basin__local_drain_direction`, `land_surface__slope`,
// `river_location__mask`).  It is not part of the per-step path.
//
// Flow directions use the PCRaster LDD convention of the reference (Wflow/src/utils.jl:3-13):
//   code 1..9 -> (dx, dy) = (-1,-1) (0,-1) (1,-1) (-1,0) (0,0) (1,0) (-1,1) (0,1) (1,1)
// with 5 the pit code.
#include <cmath>
#include <cstdint>
#include <queue>
#include <vector>

namespace {

const int DX[10] = {0, -1, 0, 1, -1, 0, 1, -1, 0, 1};
const int DY[10] = {0, -1, -1, -1, 0, 0, 0, 1, 1, 1};
const int ORDER[8] = {1, 2, 3, 4, 6, 7, 8, 9};

struct Item {
    double z;
    int64_t seq;  // insertion sequence: breaks elevation ties first-in first-out
    int64_t idx;
    bool operator>(const Item& o) const { return z > o.z || (z == o.z && seq > o.seq); }
};

}  // namespace

extern "C" {

// Priority-Flood + epsilon (Barnes, Lehman & Mulla 2014, "Priority-Flood: An optimal
// depression-filling and watershed-labeling algorithm", Algorithm 3). Every cell that is
// not on the raster edge ends up strictly higher than the cell it was reached from, so a
// steepest-descent D8 map of the result has no interior pits and no cycles.
int synth_fill_depressions(const double* dem_in, int64_t nx, int64_t ny, double* dem_out) {
    if (nx < 1 || ny < 1) return -1;
    const int64_t n = nx * ny;
    std::vector<uint8_t> closed(n, 0);
    for (int64_t i = 0; i < n; ++i) dem_out[i] = dem_in[i];
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> open;
    std::queue<int64_t> pit;
    int64_t seq = 0;
    for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x)
            if (x == 0 || y == 0 || x == nx - 1 || y == ny - 1) {
                int64_t i = x + y * nx;
                closed[i] = 1;
                open.push(Item{dem_out[i], seq++, i});
            }
    while (!open.empty() || !pit.empty()) {
        int64_t c;
        if (!pit.empty()) {
            c = pit.front();
            pit.pop();
        } else {
            c = open.top().idx;
            open.pop();
        }
        const int64_t cx = c % nx, cy = c / nx;
        for (int k = 0; k < 8; ++k) {
            const int d = ORDER[k];
            const int64_t x = cx + DX[d], y = cy + DY[d];
            if (x < 0 || y < 0 || x >= nx || y >= ny) continue;
            const int64_t j = x + y * nx;
            if (closed[j]) continue;
            closed[j] = 1;
            if (dem_out[j] <= dem_out[c]) {
                dem_out[j] = std::nextafter(dem_out[c], INFINITY);
                pit.push(j);
            } else {
                open.push(Item{dem_out[j], seq++, j});
            }
        }
    }
    return 0;
}

// Steepest-descent D8: drop / distance over the in-grid neighbours in LDD code order
// 1,2,3,4,6,7,8,9; the first maximum wins ties; no neighbour with a positive drop -> pit (5).
int synth_d8(const double* dem, int64_t nx, int64_t ny, double cell, uint8_t* ldd) {
    if (nx < 1 || ny < 1) return -1;
    const double diag = cell * std::sqrt(2.0);
    for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x) {
            const int64_t i = x + y * nx;
            double best = 0.0;
            int bestd = 5;
            for (int k = 0; k < 8; ++k) {
                const int d = ORDER[k];
                const int64_t xx = x + DX[d], yy = y + DY[d];
                if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) continue;
                const double drop = dem[i] - dem[xx + yy * nx];
                if (drop <= 0.0) continue;
                const double s = drop / ((DX[d] != 0 && DY[d] != 0) ? diag : cell);
                if (s > best) {
                    best = s;
                    bestd = d;
                }
            }
            ldd[i] = (uint8_t)bestd;
        }
    return 0;
}

// Number of cells draining through each cell (itself included), and the topological depth
// (longest upstream path, in cells) of each cell. Returns the maximum depth, or -1 when the
// LDD map leaves the grid or contains a cycle.
int64_t synth_accumulate(const uint8_t* ldd, int64_t nx, int64_t ny, int64_t* area,
                         int32_t* depth) {
    const int64_t n = nx * ny;
    std::vector<int64_t> down(n, -1);
    std::vector<int32_t> indeg(n, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int d = ldd[i];
        if (d < 1 || d > 9) return -1;
        if (d == 5) continue;
        const int64_t x = i % nx + DX[d], y = i / nx + DY[d];
        if (x < 0 || y < 0 || x >= nx || y >= ny) return -1;
        down[i] = x + y * nx;
        indeg[down[i]]++;
    }
    std::vector<int64_t> stack;
    stack.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
        area[i] = 1;
        depth[i] = 0;
        if (indeg[i] == 0) stack.push_back(i);
    }
    int64_t seen = 0, maxdepth = 0;
    while (!stack.empty()) {
        const int64_t v = stack.back();
        stack.pop_back();
        ++seen;
        if (depth[v] > maxdepth) maxdepth = depth[v];
        const int64_t w = down[v];
        if (w < 0) continue;
        area[w] += area[v];
        if (depth[v] + 1 > depth[w]) depth[w] = depth[v] + 1;
        if (--indeg[w] == 0) stack.push_back(w);
    }
    return seen == n ? maxdepth : -1;
}

}  // extern "C"
