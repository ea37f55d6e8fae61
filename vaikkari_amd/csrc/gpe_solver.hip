// gpe_solver.hip -- the solver object behind the C ABI (include/vaikkari_gpe.h):
// device buffers, kernel selection and the time-stepping loop that replaces
// the reference's host loop of update<<<>>> launches
// (koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu:406-486).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "gpe_internal.hpp"
#include "gpe_fused.hpp"
#include "gpe_fused_row.hpp"
#include "gpe_hip_check.hpp"
#include "gpe_kernels.hpp"
#include "gpe_solver.hpp"

using namespace gpe;

namespace gpe {

Layout make_layout(uint32_t nx, uint32_t ny, uint32_t nz, size_t elem_bytes) {
    Layout L{};
    L.nx = (int)nx;
    L.ny = (int)ny;
    L.nz = (int)nz;
    const long long line = (long long)(128 / elem_bytes);  // elements per 128-byte line
    L.line = (int)line;
    // x = nx and, in the row before, x = -1 share the row's last padding line
    L.xp = (((long long)nx + 1 + line - 1) / line) * line;
    L.sy = 12LL * L.xp;
    L.sz = (long long)(ny + 2) * L.sy;
    // one line in front of the halo row (j, k) = (-1, -1) holds its x = -1
    L.origin = L.sz + L.sy + line;
    L.elems = (long long)(nz + 2) * L.sz + line;
    return L;
}

Trap make_trap(const double origin[3], double b, double ratio) {
    Trap t{};
    for (int a = 0; a < 3; a++) t.p0[a] = origin[a];
    t.b = b;
    t.bw = std::sqrt(8.0);
    t.rr = ratio * ratio;
    const double sq18 = 1.0 / std::sqrt(8.0);
    for (int l = 0; l < 12; l++)
        for (int a = 0; a < 3; a++) t.bpos[l][a] = sq18 * kClassPos[l][a];
    for (int n = 0; n < 4; n++) t.q[n] = sq18 * (2 * n + 1);
    return t;
}

}  // namespace gpe

// ---------------------------------------------------------------------
// kernel dispatch
// ---------------------------------------------------------------------
namespace {

constexpr int kTY = 4;  // waves per block of the observables kernel

// Tile shapes selectable per solver (gpe_config.tile; the autotuner sweeps
// them).  ty = waves (rows of 64 cells) per block, cz = layers per z-march
// block, remap = block order over the XCDs (see block_coords: 0 plain,
// 1 contiguous eighths, 2 y-bands).
struct TileShape {
    int ty, cz, remap, nt;  // nt: streaming (non-temporal) access to the updated field (real time: its
                            // read-modify-write; imaginary time: its store)
    int row;                // fused family: one independent wave per cell row (k_fused_row)
    int node;               // gather family: 1 one thread per node (k_node, ty waves = ty classes),
                            // 2 one thread per cell and group of 4 classes (k_quad, ty (group, rows) waves)
};
// gather tiles 15-18: one thread per node, 12 or 4 classes per block; 19-24:
// one thread per group of 4 classes (23, 24: y-bands per XCD)
constexpr TileShape kGatherTiles[] = {{4, 1, 0, 1}, {4, 1, 0, 0}, {4, 1, 1, 0}, {1, 1, 1, 0}, {2, 1, 1, 0},
                                      {8, 1, 1, 0}, {4, 1, 1, 1}, {2, 1, 0, 1}, {8, 1, 0, 1}, {1, 1, 0, 1},
                                      {4, 1, 2, 1}, {2, 1, 2, 1}, {8, 1, 2, 1}, {1, 1, 2, 1},
                                      {12, 1, 0, 1, 0, 1}, {12, 1, 0, 0, 0, 1}, {4, 1, 0, 1, 0, 1},
                                      {4, 1, 1, 0, 0, 1}, {3, 1, 1, 0, 0, 2}, {6, 1, 1, 0, 0, 2},
                                      {12, 1, 1, 0, 0, 2}, {6, 1, 0, 1, 0, 2}, {3, 1, 2, 1, 0, 2},
                                      {6, 1, 2, 1, 0, 2}};
constexpr TileShape kZmarchTiles[] = {{4, 8, 1, 0}, {4, 8, 0, 1}, {1, 8, 1, 0}, {2, 8, 1, 1}, {8, 8, 1, 1},
                                      {4, 4, 1, 1}, {4, 16, 1, 1}, {2, 16, 1, 0}, {1, 16, 1, 1}, {2, 4, 1, 1},
                                      {4, 4, 2, 1}, {8, 8, 2, 1}};
// fused: ty = waves per block (footprint rows, interior ty-2), cz = layers per block
// fused, row design (row = 1): ty = waves (rows) per block, cz = layers per z-chunk
constexpr TileShape kFusedTiles[] = {{8, 16, 1, 0}, {4, 16, 1, 0}, {4, 32, 1, 0}, {8, 32, 1, 0}, {8, 8, 1, 0},
                                     {16, 16, 1, 0}, {4, 8, 1, 0}, {8, 16, 0, 0}, {6, 16, 1, 0}, {6, 32, 1, 0},
                                     {2, 32, 1, 0, 1}, {1, 32, 1, 0, 1}, {4, 32, 1, 0, 1}, {2, 16, 1, 0, 1},
                                     {1, 16, 1, 0, 1}, {2, 0, 1, 0, 1}, {1, 0, 1, 0, 1}};
constexpr int kNumGatherTiles = sizeof(kGatherTiles) / sizeof(kGatherTiles[0]);
constexpr int kNumZmarchTiles = sizeof(kZmarchTiles) / sizeof(kZmarchTiles[0]);
constexpr int kNumFusedTiles = sizeof(kFusedTiles) / sizeof(kFusedTiles[0]);

int num_tiles(int kernel) {
    return kernel == GPE_KERNEL_ZMARCH ? kNumZmarchTiles : kernel == GPE_KERNEL_FUSED ? kNumFusedTiles : kNumGatherTiles;
}

// Block rows per layer of a gather tile of ty waves (as grid_of computes them).
int gather_rows(int ty, const Layout &L) {
    const int rows = ty * (64 >> row_lanes_log2(L.nx));
    return (L.ny + rows - 1) / rows;
}

// Tile index used when the caller asks for tile 0, by family, precision,
// mode and grid.
// Fused family: the one-wave-per-row tiles measured fastest, {2 waves,
// 32 layers} in FP64 and {2 waves, 16 layers} in FP32.  Z-march: tile 1.
// Gather family, a rule fitted to the gather tiles timed on MI355X (tiles
// 1-14 over 24^3 to 256^3 cells, 87x87x302, 100x60x140 and 300x300x200,
// tiles 15-24 over 16^3 to 256^3 and 87x87x302; both precisions, both modes;
// against tiles 1-14 its one-thread-per-cell part, cell_default_tile, ran at
// 99 % of the fastest on geometric mean, 92 % at worst):
//  - real time with both time levels in 256 MiB (the Infinity Cache keeps
//    them between the two passes): a cached read-modify-write in contiguous
//    eighths, in FP64 one thread per node with 4 classes per block (tile 18:
//    +4 to +15 % over one thread per cell at 56^3-80^3), in FP32 one thread
//    per cell and group of 4 classes, 3 waves per block (tile 19: +2 to +14 %
//    at 24^3-96^3); on grids of at most 1024 one-wave blocks whose rows
//    split into y-bands (32^3: latency-bound), y-bands instead: one thread per
//    cell and group of 4 classes, 3 waves per block (tile 23; 32^3: +6 % FP64,
//    +10 % FP32 over {1 wave, y-bands} (tile 14), itself 1.5x tile 4);
//  - imaginary time with both time levels in 256 MiB: cached stores in
//    contiguous eighths, since the Euler step's output is the next step's
//    input and a streaming store sends it to HBM; FP64 from 256 blocks of 4
//    waves, one thread per group of 4 classes, 3 waves per block (tile 19);
//    FP32 from 700, one thread per cell, 8 waves per block (tile 6)
//    (48^3-96^3: +5 to +30 % over streaming stores);
//  - FP64 real time beyond the cache, when the wave rows split into y-bands
//    of whole 6-wave blocks: one thread per cell and group of 4 classes with
//    streaming read-modify-write in y-bands per XCD (tile 24): +2 to +6 %
//    over one thread per cell at 96^3-256^3 (in plain order, tile 22, the
//    class groups are level with tile 7 where no y-bands fit: -4 % at 112^3,
//    +1.5 % at 87x87x302);
//  - otherwise y-bands per XCD with streaming read-modify-write, the most
//    waves per block (up to 8 in FP64 imaginary time, 4 else: tiles 13, 11,
//    12, 14 for 8, 4, 2, 1) whose block rows divide by the 8 XCDs; when none
//    does, {4 waves, contiguous eighths, streaming} (tile 7).
// At 128^3 cells this is tile 24 in FP64 (13 in imaginary time), tile 11 in
// FP32.
struct GridFacts {
    bool cached;     // both time levels fit 256 MiB
    long long gx;    // wave columns per row
    bool tiny_band;  // one-wave blocks split into y-bands, at most 1024 of them
    long long b4;    // blocks of 4 waves per launch
};
GridFacts grid_facts(uint32_t precision, const Layout &L) {
    GridFacts f;
    f.cached = 2ull * (unsigned long long)L.elems * (precision == GPE_FP64 ? 16u : 8u) <= (256ull << 20);
    f.gx = (L.nx + (1 << row_lanes_log2(L.nx)) - 1) >> row_lanes_log2(L.nx);
    const int gy1 = gather_rows(1, L);
    f.tiny_band = gy1 % 8 == 0 && f.gx * gy1 * L.nz <= 1024;
    f.b4 = f.gx * gather_rows(4, L) * L.nz;
    return f;
}

// The default among the one-thread-per-cell gather tiles (1-14).
int cell_default_tile(uint32_t precision, int mode, const Layout &L) {
    const GridFacts f = grid_facts(precision, L);
    if (mode == MODE_REAL && f.cached) return f.tiny_band ? 13 : 3;
    if (mode == MODE_IMAG && f.cached && f.b4 >= 700) return 5;
    const int wmax = precision == GPE_FP64 && mode == MODE_IMAG ? 8 : 4;
    static constexpr int kWaves[4] = {8, 4, 2, 1}, kBandTile[4] = {12, 10, 11, 13};
    for (int i = 0; i < 4; i++)
        if (kWaves[i] <= wmax && gather_rows(kWaves[i], L) % 8 == 0) return kBandTile[i];
    return 6;
}

int default_tile(int kernel, uint32_t precision, int mode, const Layout &L) {
    if (kernel == GPE_KERNEL_FUSED) return precision == GPE_FP64 ? 10 : 13;
    if (kernel != GPE_KERNEL_GATHER) return 0;
    const GridFacts f = grid_facts(precision, L);
    // tiny grids: class groups in y-bands (3-wave blocks always split the bands
    // of wave rows evenly, as their count divides by 8)
    if (mode == MODE_REAL && f.cached) return f.tiny_band ? 22 : precision == GPE_FP64 ? 17 : 18;
    if (mode == MODE_IMAG && f.cached && precision == GPE_FP64 && f.b4 >= 256) return 18;
    if (mode == MODE_REAL && precision == GPE_FP64 && !f.cached) {
        const int rows = gather_rows(1, L);  // wave rows per layer
        if (rows % 8 == 0 && (3 * f.gx * (rows / 8)) % 6 == 0) return 23;
    }
    return cell_default_tile(precision, mode, L);
}

TileShape tile_shape(int kernel, int tile) {
    return kernel == GPE_KERNEL_ZMARCH ? kZmarchTiles[tile]
           : kernel == GPE_KERNEL_FUSED ? kFusedTiles[tile]
                                        : kGatherTiles[tile];
}

// A fused tile of 16 rows holds its FP64 exchange slots in more than the
// 160 KiB of LDS, and the one-wave-per-row fused tiles address a field with
// 32-bit byte offsets: those (family, tile, precision, size) are refused up
// front.
bool tile_supported(int kernel, int tile, uint32_t precision, const Layout &L) {
    if (kernel != GPE_KERNEL_FUSED) return true;
    const TileShape t = tile_shape(kernel, tile);
    if (t.row) return (unsigned long long)L.elems * (precision == GPE_FP64 ? 16u : 8u) < (1ull << 31);
    return !(precision == GPE_FP64 && t.ty == 16);
}

FusedGrid make_fused_grid(const Layout &L, const TileShape &t) {
    FusedGrid G;
    G.ntx = (L.nx + 61) / 62;
    G.w = (L.nx + G.ntx - 1) / G.ntx;
    G.nty = (L.ny + t.ty - 3) / (t.ty - 2);
    G.h = (L.ny + G.nty - 1) / G.nty;
    G.ntz = (L.nz + t.cz - 1) / t.cz;
    G.cz = (L.nz + G.ntz - 1) / G.ntz;
    G.remap = t.remap;
    return G;
}

template <typename T, int FORM, int POT, int TY>
void launch_fused_t(cplx_t<T> *eo, cplx_t<T> *oo, const cplx_t<T> *ei, const cplx_t<T> *oi, const T *pot,
                    const Layout &L, const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, const FusedGrid &G,
                    hipStream_t st) {
    hipLaunchKernelGGL((k_fused<T, FORM, POT, TY>), dim3(G.ntx * G.nty * G.ntz), dim3(64, TY), 0, st, eo, oo, ei, oi,
                       pot, L, H, trap, dt, G);
}

// Grid of the row-fused kernel: x-strips of 62 interior columns (the last
// one partial); z-chunks of cz layers, or (cz = 0) as many layers as make one
// round of waves fill every wave slot of the chip, since each wave marches
// its chunk from start to end and a second, partial round would leave most
// slots idle.
template <typename T, int FORM, int POT, int TY, int CZ>
hipError_t launch_fused_row_t(cplx_t<T> *eo, cplx_t<T> *oo, const cplx_t<T> *ei, const cplx_t<T> *oi, const T *pot,
                              const Layout &L, const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, int cz,
                              int remap, hipStream_t st) {
    const auto kern = k_fused_row<T, FORM, POT, TY, CZ>;
    RowGrid G;
    G.ntx = (L.nx + 61) / 62;
    G.w = 62;  // full strips measured faster than strips of nx / ntx columns
    G.nty = (L.ny + TY - 1) / TY;
    if (cz <= 0) {
        static int blocks_per_cu = -1;  // per instantiation: a property of the compiled kernel
        if (blocks_per_cu < 0) {
            int nb = 0;
            const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 64 * TY, 0);
            if (e != hipSuccess) return e;
            blocks_per_cu = nb > 0 ? nb : 1;
        }
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        const long long slots = (long long)blocks_per_cu * cus;  // blocks resident at once
        const long long cols = (long long)G.ntx * G.nty;
        const int ntz = (int)std::max(1LL, std::min<long long>(L.nz, slots / cols));
        cz = (L.nz + ntz - 1) / ntz;
    }
    G.cz = cz;
    G.ntz = (L.nz + cz - 1) / cz;
    G.remap = remap;
    hipLaunchKernelGGL(kern, dim3(G.ntx * G.nty * G.ntz), dim3(64, TY), 0, st, eo, oo, ei, oi, pot, L, H, trap, dt,
                       G);
    return hipGetLastError();
}

template <typename T, int FORM, int POT>
hipError_t launch_fused(int tile, cplx_t<T> *eo, cplx_t<T> *oo, const cplx_t<T> *ei, const cplx_t<T> *oi,
                        const T *pot, const Layout &L, const Hamiltonian<T, FORM> &H, const Trap &trap, double dt,
                        hipStream_t st) {
    const TileShape t = tile_shape(GPE_KERNEL_FUSED, tile);
    if (t.row) {
#define GPE_ROW(TYV, CZV) \
    return launch_fused_row_t<T, FORM, POT, TYV, CZV>(eo, oo, ei, oi, pot, L, H, trap, dt, CZV, t.remap, st)
        switch (t.ty * 1000 + t.cz) {
            case 1000: GPE_ROW(1, 0);
            case 2000: GPE_ROW(2, 0);
            case 1016: GPE_ROW(1, 16);
            case 2016: GPE_ROW(2, 16);
            case 1032: GPE_ROW(1, 32);
            case 2032: GPE_ROW(2, 32);
            case 4032: GPE_ROW(4, 32);
            default: return hipErrorInvalidValue;
        }
#undef GPE_ROW
    }
    const FusedGrid G = make_fused_grid(L, t);
    switch (t.ty) {
        case 4: launch_fused_t<T, FORM, POT, 4>(eo, oo, ei, oi, pot, L, H, trap, dt, G, st); break;
        case 6: launch_fused_t<T, FORM, POT, 6>(eo, oo, ei, oi, pot, L, H, trap, dt, G, st); break;
        case 8: launch_fused_t<T, FORM, POT, 8>(eo, oo, ei, oi, pot, L, H, trap, dt, G, st); break;
        case 16:
            // 16 rows of FP64 exchange slots exceed the 160 KiB of LDS
            if constexpr (sizeof(T) == 4) launch_fused_t<T, FORM, POT, 16>(eo, oo, ei, oi, pot, L, H, trap, dt, G, st);
            else return hipErrorInvalidConfiguration;
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

Grid3 grid_of(int kernel, int tile, const Layout &L) {
    const TileShape t = tile_shape(kernel, tile);
    Grid3 G{};
    G.lsh = row_lanes_log2(L.nx);
    if (t.node) {  // k_node, k_quad: gx x gy wave rows per layer, 12 classes or 3 groups each
        G.gx = (L.nx + (1 << G.lsh) - 1) >> G.lsh;
        G.gy = (L.ny + (64 >> G.lsh) - 1) / (64 >> G.lsh);
        G.gz = L.nz;
        G.remap = t.remap;
        // y-bands (k_quad only): whole blocks per band and layer
        if (G.remap == 2 && (t.node != 2 || G.gy % 8 != 0 || (3 * G.gx * (G.gy / 8)) % t.ty != 0)) G.remap = 0;
        return G;
    }
    const int lw = 1 << G.lsh, rows = t.ty * (64 >> G.lsh);
    G.gx = (L.nx + lw - 1) / lw;
    G.gy = (L.ny + rows - 1) / rows;
    G.gz = kernel == GPE_KERNEL_ZMARCH ? (L.nz + t.cz - 1) / t.cz : L.nz;
    G.remap = t.remap;
    if (G.remap == 2 && G.gy % 8 != 0) G.remap = 0;  // y-bands need gy divisible by the 8 XCDs
    return G;
}

template <typename T, int FORM, int POT, int MODE, int TY, bool NT = false>
void launch_gather(cplx_t<T> *next, const cplx_t<T> *prev, const T *pot, const Layout &L,
                   const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, const double *scale, double *partials,
                   const Grid3 &G, hipStream_t st) {
    hipLaunchKernelGGL((k_gather<T, FORM, POT, MODE, TY, NT>), dim3(G.gx * G.gy * G.gz), dim3(64, TY), 0, st, next,
                       prev, pot, L, H, trap, dt, scale, partials, G);
}

// Blocks of one launch of (kernel, tile) on grid G.
long long grid_blocks(const TileShape &t, const Grid3 &G) {
    const long long cells = (long long)G.gx * G.gy * G.gz;
    const long long per = t.node == 1 ? 12 : t.node == 2 ? 3 : 0;  // waves per wave row of cells
    return t.node ? (per * cells + t.ty - 1) / t.ty : cells;
}

template <typename T, int FORM, int POT, int MODE, int TY, bool NT = false>
void launch_node(cplx_t<T> *next, const cplx_t<T> *prev, const T *pot, const Layout &L,
                 const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, const double *scale, double *partials,
                 const Grid3 &G, hipStream_t st) {
    const long long nb = grid_blocks(TileShape{TY, 1, 0, 0, 0, 1}, G);
    hipLaunchKernelGGL((k_node<T, FORM, POT, MODE, TY, NT>), dim3((unsigned)nb), dim3(64, TY), 0, st, next, prev, pot,
                       L, H, trap, dt, scale, partials, G);
}

template <typename T, int FORM, int POT, int MODE, int TY, bool NT = false>
void launch_quad(cplx_t<T> *next, const cplx_t<T> *prev, const T *pot, const Layout &L,
                 const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, const double *scale, double *partials,
                 const Grid3 &G, hipStream_t st) {
    const long long nb = grid_blocks(TileShape{TY, 1, 0, 0, 0, 2}, G);
    hipLaunchKernelGGL((k_quad<T, FORM, POT, MODE, TY, NT>), dim3((unsigned)nb), dim3(64, TY), 0, st, next, prev, pot,
                       L, H, trap, dt, scale, partials, G);
}

template <typename T, int FORM, int POT, int MODE, int TY, int CZ, bool NT = false>
void launch_zmarch(cplx_t<T> *next, const cplx_t<T> *prev, const T *pot, const Layout &L,
                   const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, const double *scale, double *partials,
                   const Grid3 &G, hipStream_t st) {
    hipLaunchKernelGGL((k_zmarch<T, FORM, POT, MODE, TY, CZ, NT>), dim3(G.gx * G.gy * G.gz), dim3(64, TY), 0, st,
                       next, prev, pot, L, H, trap, dt, scale, partials, G);
}

template <typename T, int FORM, int POT, int MODE>
hipError_t launch_update(int kernel, int tile, cplx_t<T> *next, const cplx_t<T> *prev, const T *pot, const Layout &L,
                         const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, const double *scale,
                         double *partials, hipStream_t st, bool reverse = false) {
    Grid3 G = grid_of(kernel, tile, L);
    G.rev = reverse ? 1 : 0;
    const TileShape t = tile_shape(kernel, tile);
#define GPE_ARGS next, prev, pot, L, H, trap, dt, scale, partials, G, st
    if (kernel == GPE_KERNEL_ZMARCH) {
        switch (t.ty * 1000 + t.cz * 10 + (t.nt ? 1 : 0)) {
            case 1080: launch_zmarch<T, FORM, POT, MODE, 1, 8>(GPE_ARGS); break;
            case 1161: launch_zmarch<T, FORM, POT, MODE, 1, 16, true>(GPE_ARGS); break;
            case 1160: launch_zmarch<T, FORM, POT, MODE, 1, 16>(GPE_ARGS); break;
            case 2041: launch_zmarch<T, FORM, POT, MODE, 2, 4, true>(GPE_ARGS); break;
            case 2040: launch_zmarch<T, FORM, POT, MODE, 2, 4>(GPE_ARGS); break;
            case 2081: launch_zmarch<T, FORM, POT, MODE, 2, 8, true>(GPE_ARGS); break;
            case 2080: launch_zmarch<T, FORM, POT, MODE, 2, 8>(GPE_ARGS); break;
            case 2160: launch_zmarch<T, FORM, POT, MODE, 2, 16>(GPE_ARGS); break;
            case 4041: launch_zmarch<T, FORM, POT, MODE, 4, 4, true>(GPE_ARGS); break;
            case 4040: launch_zmarch<T, FORM, POT, MODE, 4, 4>(GPE_ARGS); break;
            case 4080: launch_zmarch<T, FORM, POT, MODE, 4, 8>(GPE_ARGS); break;
            case 4081: launch_zmarch<T, FORM, POT, MODE, 4, 8, true>(GPE_ARGS); break;
            case 4161: launch_zmarch<T, FORM, POT, MODE, 4, 16, true>(GPE_ARGS); break;
            case 4160: launch_zmarch<T, FORM, POT, MODE, 4, 16>(GPE_ARGS); break;
            case 8081: launch_zmarch<T, FORM, POT, MODE, 8, 8, true>(GPE_ARGS); break;
            case 8080: launch_zmarch<T, FORM, POT, MODE, 8, 8>(GPE_ARGS); break;
            default: return hipErrorInvalidValue;
        }
    } else if (t.node == 2) {
        switch (t.ty * 10 + t.nt) {
            case 30: launch_quad<T, FORM, POT, MODE, 3>(GPE_ARGS); break;
            case 60: launch_quad<T, FORM, POT, MODE, 6>(GPE_ARGS); break;
            case 61: launch_quad<T, FORM, POT, MODE, 6, true>(GPE_ARGS); break;
            case 31: launch_quad<T, FORM, POT, MODE, 3, true>(GPE_ARGS); break;
            case 120: launch_quad<T, FORM, POT, MODE, 12>(GPE_ARGS); break;
            default: return hipErrorInvalidValue;
        }
    } else if (t.node) {
        switch (t.ty * 10 + t.nt) {
            case 40: launch_node<T, FORM, POT, MODE, 4>(GPE_ARGS); break;
            case 41: launch_node<T, FORM, POT, MODE, 4, true>(GPE_ARGS); break;
            case 120: launch_node<T, FORM, POT, MODE, 12>(GPE_ARGS); break;
            case 121: launch_node<T, FORM, POT, MODE, 12, true>(GPE_ARGS); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (t.ty * 10 + t.nt) {
            case 10: launch_gather<T, FORM, POT, MODE, 1>(GPE_ARGS); break;
            case 20: launch_gather<T, FORM, POT, MODE, 2>(GPE_ARGS); break;
            case 40: launch_gather<T, FORM, POT, MODE, 4>(GPE_ARGS); break;
            case 80: launch_gather<T, FORM, POT, MODE, 8>(GPE_ARGS); break;
            case 11: launch_gather<T, FORM, POT, MODE, 1, true>(GPE_ARGS); break;
            case 21: launch_gather<T, FORM, POT, MODE, 2, true>(GPE_ARGS); break;
            case 41: launch_gather<T, FORM, POT, MODE, 4, true>(GPE_ARGS); break;
            case 81: launch_gather<T, FORM, POT, MODE, 8, true>(GPE_ARGS); break;
            default: return hipErrorInvalidValue;
        }
    }
#undef GPE_ARGS
    return hipGetLastError();
}

long long update_blocks(int kernel, int tile, const Layout &L) {
    return grid_blocks(tile_shape(kernel, tile), grid_of(kernel, tile, L));
}

long long max_update_blocks(const Layout &L) {
    long long m = 0;
    for (int t = 0; t < kNumGatherTiles; t++) m = std::max(m, update_blocks(GPE_KERNEL_GATHER, t, L));
    for (int t = 0; t < kNumZmarchTiles; t++) m = std::max(m, update_blocks(GPE_KERNEL_ZMARCH, t, L));
    return m;
}

long long obs_blocks(const Layout &L) { return (long long)((L.nx + 63) / 64) * ((L.ny + kTY - 1) / kTY) * L.nz; }

}  // namespace

// ---------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------
gpe_solver::~gpe_solver() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto *p : field)
        if (p) (void)hipFree(p);
    if (pot) (void)hipFree(pot);
    if (scal) (void)hipFree(scal);
    if (partials) (void)hipFree(partials);
    if (out_buf) (void)hipFree(out_buf);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
}

size_t gpe_solver::elem_bytes() const { return cfg.precision == GPE_FP64 ? 16 : 8; }
size_t gpe_solver::nodes() const { return (size_t)cfg.nx * cfg.ny * cfg.nz * GPE_CLASSES; }

static gpe_status validate(const gpe_config *c) {
    if (!c) {
        set_error("gpe_create: null config");
        return GPE_EINVAL;
    }
    if (c->nx == 0 || c->ny == 0 || c->nz == 0 || c->nx > 65536 || c->ny > 65536 || c->nz > 65536) {
        set_error("gpe_create: grid %u x %u x %u out of range", c->nx, c->ny, c->nz);
        return GPE_EINVAL;
    }
    if (c->precision > GPE_FP32_BULK || c->mode > GPE_IMAG_TIME || c->potential > GPE_POT_TRAP ||
        c->kernel > GPE_KERNEL_FUSED) {
        set_error("gpe_create: bad enum value");
        return GPE_EINVAL;
    }
    if (!(c->block_scale > 0) || !(c->dt > 0) || !std::isfinite(c->g) || !std::isfinite(c->trap_ratio)) {
        set_error("gpe_create: block_scale and dt must be positive, g and trap_ratio finite");
        return GPE_EINVAL;
    }
    if (c->potential == GPE_POT_TRAP) {
        const uint32_t dims[3] = {c->nx, c->ny, c->nz};
        const double vmax = trap_vmax(dims, c->origin, c->block_scale, c->trap_ratio);
        if (!linear_step_stable(c->dt, c->block_scale, vmax)) {
            set_error("gpe_create: dt = %g violates dt*(6/b^2 + max V) < 2 (max V = %g)", c->dt, vmax);
            return GPE_EUNSTABLE;
        }
    }
    return GPE_OK;
}

gpe_status gpe_solver::init(const gpe_config &c) {
    cfg = c;
    L = make_layout(c.nx, c.ny, c.nz, elem_bytes());
    trap = make_trap(c.origin, c.block_scale, c.trap_ratio);
    // AUTO: the one-thread-per-cell gather pass, the fastest family measured on
    // MI355X for both precisions (the z-march and fused families stay
    // selectable).
    kernel = c.kernel == GPE_KERNEL_AUTO ? GPE_KERNEL_GATHER : c.kernel;
    if ((int)c.tile > num_tiles(kernel)) {
        set_error("gpe_create: tile %u out of range for this kernel family", c.tile);
        return GPE_EINVAL;
    }
    tile = c.tile == 0 ? default_tile(kernel, c.precision, MODE_REAL, L) : (int)c.tile - 1;
    tile_default = c.tile == 0;
    if (!tile_supported(kernel, tile, c.precision, L)) {
        set_error("gpe_create: fused tile %u does not fit LDS in FP64, or is a row tile on a field of 2 GiB or more",
                  c.tile);
        return GPE_EINVAL;
    }
    GPE_HIP_TRY(hipGetDevice(&device));
    GPE_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    GPE_HIP_TRY(hipEventCreate(&ev0));
    GPE_HIP_TRY(hipEventCreate(&ev1));
    for (int f = 0; f < 2; f++) {
        gpe_status st = alloc_field(f);
        if (st != GPE_OK) return st;
    }
    if (kernel == GPE_KERNEL_FUSED) {
        for (int f = 2; f < 4; f++) {
            gpe_status st = alloc_field(f);
            if (st != GPE_OK) return st;
        }
    }
    npartials = std::max(max_update_blocks(L), obs_blocks(L));
    GPE_HIP_TRY(hipMalloc(&partials, (size_t)npartials * 4 * sizeof(double)));
    GPE_HIP_TRY(hipMalloc(&scal, 8 * sizeof(double)));
    const double init_scal[8] = {1.0, 1.0, 0, 0, 0, 0, 0, 0};
    GPE_HIP_TRY(hipMemcpyAsync(scal, init_scal, sizeof(init_scal), hipMemcpyHostToDevice, stream));
    GPE_HIP_TRY(hipStreamSynchronize(stream));
    set_coefs();
    return GPE_OK;
}

gpe_status gpe_solver::alloc_field(int f) {
    if (field[f]) return GPE_OK;
    const size_t fbytes = (size_t)L.elems * elem_bytes();
    if (hipMalloc(&field[f], fbytes) != hipSuccess) {
        (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
        field[f] = nullptr;
        set_error("gpe_create: cannot allocate %zu bytes of device memory", fbytes);
        return GPE_ENOMEM;
    }
    // the zero halo of every field is never written afterwards
    GPE_HIP_TRY(hipMemsetAsync(field[f], 0, fbytes, stream));
    GPE_HIP_TRY(hipStreamSynchronize(stream));
    return GPE_OK;
}

void gpe_solver::set_coefs() {
    const double b = cfg.block_scale;
    const double lapfac = -0.5 * kHodge / (b * b);
    const double lapfac0 = (double)GPE_FACES * (-lapfac);
    co.lapfac_dt = lapfac * cfg.dt;
    co.lapfac0_dt = lapfac0 * cfg.dt;
    co.g_dt = cfg.g * cfg.dt;
    co.dt = cfg.dt;
    co.w = node_measure(b);
    co.lap = kHodge / (b * b);
    co.g = cfg.g;
}

size_t gpe_solver::device_bytes() const {
    size_t b = 0;
    for (auto *p : field)
        if (p) b += (size_t)L.elems * elem_bytes();
    if (pot) b += (size_t)L.elems * (elem_bytes() / 2);
    return b + (size_t)npartials * 4 * sizeof(double) + 8 * sizeof(double);
}

// Host [k][j][i][l] complex double <-> device class-row layout.
template <typename T>
static void pack_field(const double *h, std::vector<T> &d, const Layout &L) {
    d.assign((size_t)L.elems * 2, T(0));
    size_t n = 0;
    for (int k = 0; k < L.nz; k++)
        for (int j = 0; j < L.ny; j++)
            for (int i = 0; i < L.nx; i++)
                for (int l = 0; l < 12; l++, n++) {
                    const long long e = L.origin + (long long)k * L.sz + (long long)j * L.sy + l * L.xp + i;
                    d[2 * e] = (T)h[2 * n];
                    d[2 * e + 1] = (T)h[2 * n + 1];
                }
}
template <typename T>
static void unpack_field(const std::vector<T> &d, double *h, const Layout &L) {
    size_t n = 0;
    for (int k = 0; k < L.nz; k++)
        for (int j = 0; j < L.ny; j++)
            for (int i = 0; i < L.nx; i++)
                for (int l = 0; l < 12; l++, n++) {
                    const long long e = L.origin + (long long)k * L.sz + (long long)j * L.sy + l * L.xp + i;
                    h[2 * n] = (double)d[2 * e];
                    h[2 * n + 1] = (double)d[2 * e + 1];
                }
}

namespace gpe {

template <typename T>
gpe_status upload_field(void *dev, const double *h, const Layout &L, hipStream_t st) {
    std::vector<T> d;
    pack_field<T>(h, d, L);
    GPE_HIP_TRY(hipMemcpyAsync(dev, d.data(), d.size() * sizeof(T), hipMemcpyHostToDevice, st));
    GPE_HIP_TRY(hipStreamSynchronize(st));
    return GPE_OK;
}

template <typename T>
gpe_status download_field(const void *dev, double *h, const Layout &L, hipStream_t st) {
    std::vector<T> d((size_t)L.elems * 2);
    GPE_HIP_TRY(hipMemcpyAsync(d.data(), dev, d.size() * sizeof(T), hipMemcpyDeviceToHost, st));
    GPE_HIP_TRY(hipStreamSynchronize(st));
    unpack_field<T>(d, h, L);
    return GPE_OK;
}

template gpe_status upload_field<double>(void *, const double *, const Layout &, hipStream_t);
template gpe_status upload_field<float>(void *, const double *, const Layout &, hipStream_t);
template gpe_status download_field<double>(const void *, double *, const Layout &, hipStream_t);
template gpe_status download_field<float>(const void *, double *, const Layout &, hipStream_t);

}  // namespace gpe

template <typename T>
gpe_status gpe_solver::upload(int which, const double *h) {
    return upload_field<T>(field[which], h, L, stream);
}

template <typename T>
gpe_status gpe_solver::download(int which, double *h) {
    return download_field<T>(field[which], h, L, stream);
}

gpe_status gpe_solver::upload_potential() {
    // V*dt, multiplied on the host exactly as GPUGrossPitaevskii_BCC.cu:307.
    const size_t eb = elem_bytes() / 2;
    std::vector<double> d64;
    std::vector<float> d32;
    if (eb == 8) d64.assign((size_t)L.elems, 0.0);
    else d32.assign((size_t)L.elems, 0.0f);
    size_t n = 0;
    for (int k = 0; k < L.nz; k++)
        for (int j = 0; j < L.ny; j++)
            for (int i = 0; i < L.nx; i++)
                for (int l = 0; l < 12; l++, n++) {
                    const long long e = L.origin + (long long)k * L.sz + (long long)j * L.sy + l * L.xp + i;
                    const double v = V_host[n] * cfg.dt;
                    if (eb == 8) d64[e] = v;
                    else d32[e] = (float)v;
                }
    if (!pot && hipMalloc(&pot, (size_t)L.elems * eb) != hipSuccess) {
        (void)hipGetLastError();
        pot = nullptr;
        set_error("gpe_set_potential: cannot allocate the potential");
        return GPE_ENOMEM;
    }
    GPE_HIP_TRY(hipMemcpyAsync(pot, eb == 8 ? (void *)d64.data() : (void *)d32.data(), (size_t)L.elems * eb,
                               hipMemcpyHostToDevice, stream));
    GPE_HIP_TRY(hipStreamSynchronize(stream));
    return GPE_OK;
}

// The family and tile of the imaginary-time Euler step: the configured ones,
// the family's imaginary-time default for tile 0, and the gather family's
// default for the fused family (whose pass is a leapfrog step).
void gpe_solver::imag_variant(int &k, int &t) const {
    const bool fused = kernel == GPE_KERNEL_FUSED;
    k = fused ? GPE_KERNEL_GATHER : kernel;
    t = fused || tile_default ? default_tile(k, cfg.precision, MODE_IMAG, L) : tile;
}

template <typename T, int FORM>
gpe_status gpe_solver::enqueue_steps(uint64_t nsteps, uint64_t *launches) {
    using C = cplx_t<T>;
    Hamiltonian<T, FORM> H;
    H.lapfac = (T)co.lapfac_dt;
    H.lapfac0 = (T)co.lapfac0_dt;
    H.g = (T)co.g_dt;
    const T *p = (const T *)pot;
    const bool arr = cfg.potential == GPE_POT_ARRAY;
    uint64_t nl = 0;
    hipError_t e = hipSuccess;
    if (cfg.mode == GPE_REAL_TIME && kernel == GPE_KERNEL_FUSED) {
        int sp[2], n = 0;
        for (int f = 0; f < 4; f++)
            if (f != ie && f != io) sp[n++] = f;
        for (uint64_t s = 0; s < nsteps && e == hipSuccess; s++) {
            C *eo = (C *)field[sp[0]], *oo = (C *)field[sp[1]];
            const C *ei = (const C *)field[ie], *oi = (const C *)field[io];
            if (arr) e = launch_fused<T, FORM, POT_ARRAY>(tile, eo, oo, ei, oi, p, L, H, trap, co.dt, stream);
            else e = launch_fused<T, FORM, POT_TRAP>(tile, eo, oo, ei, oi, p, L, H, trap, co.dt, stream);
            if (e != hipSuccess) break;  // nothing was written: the state stays where it is
            std::swap(ie, sp[0]);
            std::swap(io, sp[1]);
            nl += 1;
        }
    } else if (cfg.mode == GPE_REAL_TIME) {
        C *ev = (C *)field[ie], *od = (C *)field[io];
        for (uint64_t s = 0; s < nsteps && e == hipSuccess; s++) {
            // The even update walks the tiles backwards: it starts on the layers
            // the odd update touched last, which are still in the caches.
            if (arr) {
                e = launch_update<T, FORM, POT_ARRAY, MODE_REAL>(kernel, tile, od, ev, p, L, H, trap, co.dt, nullptr, nullptr, stream);
                if (e == hipSuccess)
                    e = launch_update<T, FORM, POT_ARRAY, MODE_REAL>(kernel, tile, ev, od, p, L, H, trap, co.dt, nullptr, nullptr, stream, true);
            } else {
                e = launch_update<T, FORM, POT_TRAP, MODE_REAL>(kernel, tile, od, ev, p, L, H, trap, co.dt, nullptr, nullptr, stream);
                if (e == hipSuccess)
                    e = launch_update<T, FORM, POT_TRAP, MODE_REAL>(kernel, tile, ev, od, p, L, H, trap, co.dt, nullptr, nullptr, stream, true);
            }
            nl += 2;
        }
    } else {
        // Euler step into the other buffer; renormalisation folded into the
        // next read through scal[1] = 1/sqrt(N).  (The fused pass is a
        // leapfrog step; imaginary time runs the gather family's default.)
        int kernel, tile;
        imag_variant(kernel, tile);
        for (uint64_t s = 0; s < nsteps && e == hipSuccess; s++) {
            C *in = (C *)field[ie], *out = (C *)field[io];
            const bool rev = (steps_done + s) & 1;  // alternate directions: start where the last step ended
            if (arr)
                e = launch_update<T, FORM, POT_ARRAY, MODE_IMAG>(kernel, tile, out, in, p, L, H, trap, co.dt, scal + 1, partials, stream, rev);
            else
                e = launch_update<T, FORM, POT_TRAP, MODE_IMAG>(kernel, tile, out, in, p, L, H, trap, co.dt, scal + 1, partials, stream, rev);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(256), 0, stream, partials, update_blocks(kernel, tile, L), co.w, 1, scal);
            e = hipGetLastError();
            std::swap(ie, io);
            nl += 1;
        }
        if (e == hipSuccess && nsteps > 0) {
            // leave a normalised field behind: field[ie] *= scal[1], scal[1] = 1
            hipLaunchKernelGGL((k_scale_field<T>), dim3(1024), dim3(256), 0, stream, (C *)field[ie], L.elems, scal + 1);
            e = hipGetLastError();
            const double one = 1.0;
            if (e == hipSuccess) e = hipMemcpyAsync(scal + 1, &one, sizeof(double), hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);  // `one` lives on this stack frame
        }
    }
    if (e != hipSuccess) {
        set_error("gpe_step: launch failed: %s", hipGetErrorString(e));
        return GPE_EDEVICE;
    }
    steps_done += nsteps;
    if (launches) *launches = nl;
    return GPE_OK;
}

// One real-time step whose odd half-update (gather, `tile`, one thread per
// cell) also reduces the observables of the even field into scal[4..7]; the
// even half-update runs gather tile `even_tile`.
template <typename T, int FORM>
gpe_status gpe_solver::observe_step(int tile, int even_tile) {
    using C = cplx_t<T>;
    Hamiltonian<T, FORM> H;
    H.lapfac = (T)co.lapfac_dt;
    H.lapfac0 = (T)co.lapfac0_dt;
    H.g = (T)co.g_dt;
    const T *p = (const T *)pot;
    const bool arr = cfg.potential == GPE_POT_ARRAY;
    C *ev = (C *)field[ie], *od = (C *)field[io];
    const TileShape t = tile_shape(GPE_KERNEL_GATHER, tile);
    const Grid3 G = grid_of(GPE_KERNEL_GATHER, tile, L);
    const ObsCoefs oc{1.0 / cfg.dt, co.lap, co.g};
    const dim3 grid(G.gx * G.gy * G.gz);
#define GPE_OBS(POTK, TYV, NTV)                                                                                  \
    hipLaunchKernelGGL((k_gather_obs<T, FORM, POTK, TYV, NTV>), grid, dim3(64, TYV), 0, stream, od, ev, p, L, H, \
                       trap, co.dt, oc, partials, G)
#define GPE_OBS_TY(POTK)                                 \
    switch (t.ty * 10 + t.nt) {                          \
        case 10: GPE_OBS(POTK, 1, false); break;         \
        case 20: GPE_OBS(POTK, 2, false); break;         \
        case 40: GPE_OBS(POTK, 4, false); break;         \
        case 80: GPE_OBS(POTK, 8, false); break;         \
        case 11: GPE_OBS(POTK, 1, true); break;          \
        case 21: GPE_OBS(POTK, 2, true); break;          \
        case 41: GPE_OBS(POTK, 4, true); break;          \
        case 81: GPE_OBS(POTK, 8, true); break;          \
        default: set_error("gpe_step_observe: bad tile"); return GPE_EINVAL; \
    }
    if (arr) {
        GPE_OBS_TY(POT_ARRAY)
    } else {
        GPE_OBS_TY(POT_TRAP)
    }
#undef GPE_OBS_TY
#undef GPE_OBS
    GPE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_finalize<4>), dim3(1), dim3(256), 0, stream, partials, (long long)G.gx * G.gy * G.gz, co.w,
                       0, scal + 4);
    GPE_HIP_TRY(hipGetLastError());
    hipError_t e;
    if (arr)
        e = launch_update<T, FORM, POT_ARRAY, MODE_REAL>(GPE_KERNEL_GATHER, even_tile, ev, od, p, L, H, trap, co.dt,
                                                          nullptr, nullptr, stream, true);
    else
        e = launch_update<T, FORM, POT_TRAP, MODE_REAL>(GPE_KERNEL_GATHER, even_tile, ev, od, p, L, H, trap, co.dt,
                                                         nullptr, nullptr, stream, true);
    GPE_HIP_TRY(e);
    steps_done += 1;
    return GPE_OK;
}

gpe_status gpe_solver::step(uint64_t nsteps, uint64_t *launches) {
    if (cfg.mode == GPE_REAL_TIME && !odd_valid) {
        set_error("gpe_step: real-time stepping needs an odd field (gpe_set_state or gpe_init_odd)");
        return GPE_EINVAL;
    }
    if (cfg.potential == GPE_POT_ARRAY && !pot) {
        set_error("gpe_step: potential kind is ARRAY but no potential was set");
        return GPE_EINVAL;
    }
    if (cfg.precision == GPE_FP64) return enqueue_steps<double, FORM_REF>(nsteps, launches);
    return enqueue_steps<float, FORM_DIFF>(nsteps, launches);
}

template <typename T>
gpe_status gpe_solver::observables_t(int which, double out[4]) {
    const long long nb = obs_blocks(L);
    const dim3 grid((L.nx + 63) / 64, (L.ny + kTY - 1) / kTY, L.nz), block(64, kTY);
    const cplx_t<T> *f = (const cplx_t<T> *)field[which];
    if (cfg.potential == GPE_POT_ARRAY)
        hipLaunchKernelGGL((k_observables<T, POT_ARRAY, kTY>), grid, block, 0, stream, f, (const T *)pot, L, trap,
                           1.0 / cfg.dt, co.lap, co.g, partials);
    else
        hipLaunchKernelGGL((k_observables<T, POT_TRAP, kTY>), grid, block, 0, stream, f, (const T *)pot, L, trap,
                           1.0 / cfg.dt, co.lap, co.g, partials);
    GPE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_finalize<4>), dim3(1), dim3(256), 0, stream, partials, nb, co.w, 0, scal + 4);
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipMemcpyAsync(out, scal + 4, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
    GPE_HIP_TRY(hipStreamSynchronize(stream));
    return GPE_OK;
}

template <typename TA, typename TB>
static void launch_distance(gpe_solver *a, int fa, const gpe_solver *b, int fb) {
    const dim3 grid((a->L.nx + 63) / 64, (a->L.ny + kTY - 1) / kTY, a->L.nz), block(64, kTY);
    hipLaunchKernelGGL((k_field_distance<TA, TB, kTY>), grid, block, 0, a->stream,
                       (const cplx_t<TA> *)a->field[fa], a->L, (const cplx_t<TB> *)b->field[fb], b->L, a->partials);
}

// ---------------------------------------------------------------------
// allocation-free observables (gpe_observables_noalloc)
// ---------------------------------------------------------------------
namespace {
// Block partials of gpe_observables_noalloc, 4 doubles per block of the
// observables grid (64 x 4 cells per block and layer): grids of up to this many
// blocks, about 390^3 cells.
constexpr long long kNoAllocBlocks = 1LL << 18;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
}  // namespace

// 8 MiB per GPU, held until the process ends.  Every block of the
// observables kernel writes its own 4 partials before the finalize reads
// them, so nothing in it needs clearing between calls.  In host code the name
// below is a host-side stand-in: kernels get the address scratch_address()
// looked up on the current device.
static __device__ __align__(256) double scratch[4 * kNoAllocBlocks];

// Device address of `scratch` on the current device, looked up once per device.
static gpe_status scratch_address(double **out) {
    int dev = 0;
    GPE_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("gpe_observables_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return GPE_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        GPE_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return GPE_OK;
}

template <typename T>
static gpe_status observables_noalloc_t(const gpe_config &c, const Layout &L, const void *field, const void *pot,
                                        double *out, hipStream_t st, double *partials) {
    const double b = c.block_scale;
    const Trap trap = make_trap(c.origin, b, c.trap_ratio);
    const dim3 grid((L.nx + 63) / 64, (L.ny + kTY - 1) / kTY, L.nz), block(64, kTY);
    const cplx_t<T> *f = (const cplx_t<T> *)field;
    const double lap = kHodge / (b * b);
    if (c.potential == GPE_POT_ARRAY)
        hipLaunchKernelGGL((k_observables<T, POT_ARRAY, kTY>), grid, block, 0, st, f, (const T *)pot, L, trap,
                           1.0 / c.dt, lap, c.g, partials);
    else
        hipLaunchKernelGGL((k_observables<T, POT_TRAP, kTY>), grid, block, 0, st, f, (const T *)nullptr, L, trap,
                           1.0 / c.dt, lap, c.g, partials);
    GPE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_finalize<4>), dim3(1), dim3(256), 0, st, partials, obs_blocks(L), node_measure(b), 0, out);
    GPE_HIP_TRY(hipGetLastError());
    return GPE_OK;
}

template <typename T>
static gpe_status init_gaussian_t(gpe_solver *s, double sigma, double noise, uint64_t seed) {
    using C = cplx_t<T>;
    const double amp = std::pow(3.14159265358979323846 * sigma * sigma, -0.75);
    hipLaunchKernelGGL((k_init_gaussian<T>), dim3(1024), dim3(256), 0, s->stream, (C *)s->field[s->ie], (C *)nullptr,
                       s->L, s->trap, 1.0 / (2.0 * sigma * sigma), amp, noise, (unsigned long long)seed, 1.0, 0.0);
    GPE_HIP_TRY(hipGetLastError());
    double obs[4];
    gpe_status st = s->observables_t<T>(s->ie, obs);
    if (st != GPE_OK) return st;
    if (!(obs[0] > 0) || !std::isfinite(obs[0])) {
        set_error("gpe_init_gaussian: the Gaussian has no positive finite norm on this grid");
        return GPE_EUNSTABLE;
    }
    hipLaunchKernelGGL((k_scale_const<T>), dim3(1024), dim3(256), 0, s->stream, (C *)s->field[s->ie], s->L.elems,
                       1.0 / std::sqrt(obs[0]));
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    return GPE_OK;
}

// ---------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------
extern "C" {

gpe_status gpe_create(const gpe_config *cfg, gpe_solver **out) {
    if (!out) {
        set_error("gpe_create: null output pointer");
        return GPE_EINVAL;
    }
    *out = nullptr;
    gpe_status st = validate(cfg);
    if (st != GPE_OK) return st;
    auto *s = new (std::nothrow) gpe_solver();
    if (!s) return GPE_ENOMEM;
    st = s->init(*cfg);
    if (st != GPE_OK) {
        delete s;
        return st;
    }
    *out = s;
    return GPE_OK;
}

void gpe_destroy(gpe_solver *s) { delete s; }

gpe_status gpe_set_potential(gpe_solver *s, const double *V) {
    if (!s || !V) {
        set_error("gpe_set_potential: null argument");
        return GPE_EINVAL;
    }
    s->V_host.assign(V, V + s->nodes());
    for (double v : s->V_host)
        if (!std::isfinite(v)) {
            set_error("gpe_set_potential: non-finite potential value");
            return GPE_EINVAL;
        }
    const double vmax = *std::max_element(s->V_host.begin(), s->V_host.end());
    if (!linear_step_stable(s->cfg.dt, s->cfg.block_scale, std::max(vmax, 0.0))) {
        set_error("gpe_set_potential: dt*(6/b^2 + max V) >= 2 for this potential");
        return GPE_EUNSTABLE;
    }
    s->cfg.potential = GPE_POT_ARRAY;
    return s->upload_potential();
}

gpe_status gpe_set_state(gpe_solver *s, const double *even, const double *odd) {
    if (!s || !even) {
        set_error("gpe_set_state: null argument");
        return GPE_EINVAL;
    }
    gpe_status st;
    const bool f64 = s->cfg.precision == GPE_FP64;
    st = f64 ? s->upload<double>(s->ie, even) : s->upload<float>(s->ie, even);
    if (st != GPE_OK) return st;
    if (odd) {
        st = f64 ? s->upload<double>(s->io, odd) : s->upload<float>(s->io, odd);
        if (st != GPE_OK) return st;
        s->odd_valid = true;
    } else {
        s->odd_valid = false;
    }
    const double one = 1.0;
    GPE_HIP_TRY(hipMemcpy(s->scal + 1, &one, sizeof(double), hipMemcpyHostToDevice));
    return GPE_OK;
}

gpe_status gpe_get_state(gpe_solver *s, double *even, double *odd) {
    if (!s) {
        set_error("gpe_get_state: null solver");
        return GPE_EINVAL;
    }
    const bool f64 = s->cfg.precision == GPE_FP64;
    if (even) {
        gpe_status st = f64 ? s->download<double>(s->ie, even) : s->download<float>(s->ie, even);
        if (st != GPE_OK) return st;
    }
    if (odd) {
        gpe_status st = f64 ? s->download<double>(s->io, odd) : s->download<float>(s->io, odd);
        if (st != GPE_OK) return st;
    }
    return GPE_OK;
}

gpe_status gpe_set_mode(gpe_solver *s, uint32_t mode, double dt) {
    if (!s || mode > GPE_IMAG_TIME || !(dt > 0)) {
        set_error("gpe_set_mode: invalid argument");
        return GPE_EINVAL;
    }
    if (s->cfg.potential == GPE_POT_TRAP) {
        const uint32_t dims[3] = {s->cfg.nx, s->cfg.ny, s->cfg.nz};
        const double vmax = trap_vmax(dims, s->cfg.origin, s->cfg.block_scale, s->cfg.trap_ratio);
        if (!linear_step_stable(dt, s->cfg.block_scale, vmax)) {
            set_error("gpe_set_mode: dt = %g violates dt*(6/b^2 + max V) < 2", dt);
            return GPE_EUNSTABLE;
        }
    }
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    const bool dt_changed = dt != s->cfg.dt;
    if (mode != s->cfg.mode) s->odd_valid = false;
    s->cfg.mode = mode;
    s->cfg.dt = dt;
    s->set_coefs();
    if (dt_changed && s->pot) return s->upload_potential();
    return GPE_OK;
}

gpe_status gpe_init_odd(gpe_solver *s, double mu) {
    if (!s || !std::isfinite(mu)) {
        set_error("gpe_init_odd: invalid argument");
        return GPE_EINVAL;
    }
    // getPhase(-dt/2) = (cos(-mu dt/2), -sin(-mu dt/2)) (VortexState.cpp:198-202).
    const double ph = (-0.5 * s->cfg.dt) * mu;
    const double cr = std::cos(ph), ci = -std::sin(ph);
    if (s->cfg.precision == GPE_FP64)
        hipLaunchKernelGGL((k_phase_copy<double>), dim3(1024), dim3(256), 0, s->stream, (double2 *)s->field[s->io],
                           (const double2 *)s->field[s->ie], s->L.elems, cr, ci);
    else
        hipLaunchKernelGGL((k_phase_copy<float>), dim3(1024), dim3(256), 0, s->stream, (float2 *)s->field[s->io],
                           (const float2 *)s->field[s->ie], s->L.elems, cr, ci);
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    s->odd_valid = true;
    return GPE_OK;
}

gpe_status gpe_init_gaussian(gpe_solver *s, double sigma, double noise, uint64_t seed, double mu) {
    if (!s || !(sigma > 0) || !std::isfinite(sigma) || !std::isfinite(noise) || std::isinf(mu)) {
        set_error("gpe_init_gaussian: invalid argument");
        return GPE_EINVAL;
    }
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    const gpe_status st = s->cfg.precision == GPE_FP64 ? init_gaussian_t<double>(s, sigma, noise, seed)
                                                       : init_gaussian_t<float>(s, sigma, noise, seed);
    if (st != GPE_OK) return st;
    const double one = 1.0;
    GPE_HIP_TRY(hipMemcpy(s->scal + 1, &one, sizeof(double), hipMemcpyHostToDevice));
    s->odd_valid = false;
    if (s->cfg.mode != GPE_REAL_TIME) return GPE_OK;
    if (std::isnan(mu)) {
        double obs[4];
        const gpe_status so = gpe_observables(s, 0, obs);
        if (so != GPE_OK) return so;
        mu = (obs[1] + obs[2] + obs[3]) / obs[0];
        if (!std::isfinite(mu)) {
            set_error("gpe_init_gaussian: the state has no finite chemical potential");
            return GPE_EUNSTABLE;
        }
    }
    return gpe_init_odd(s, mu);
}

gpe_status gpe_step(gpe_solver *s, uint64_t nsteps) {
    if (!s) {
        set_error("gpe_step: null solver");
        return GPE_EINVAL;
    }
    return s->step(nsteps, nullptr);
}

gpe_status gpe_time_steps(gpe_solver *s, uint64_t nsteps, double *elapsed_ms, uint64_t *launches) {
    if (!s || !elapsed_ms) {
        set_error("gpe_time_steps: null argument");
        return GPE_EINVAL;
    }
    GPE_HIP_TRY(hipEventRecord(s->ev0, s->stream));
    gpe_status st = s->step(nsteps, launches);
    if (st != GPE_OK) return st;
    GPE_HIP_TRY(hipEventRecord(s->ev1, s->stream));
    GPE_HIP_TRY(hipEventSynchronize(s->ev1));
    float ms = 0.0f;
    GPE_HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    *elapsed_ms = ms;
    return GPE_OK;
}

gpe_status gpe_observables(gpe_solver *s, int which, double out[4]) {
    if (!s || !out || (which != 0 && which != 1)) {
        set_error("gpe_observables: invalid argument");
        return GPE_EINVAL;
    }
    const int f = which == 0 ? s->ie : s->io;
    if (s->cfg.precision == GPE_FP64) return s->observables_t<double>(f, out);
    return s->observables_t<float>(f, out);
}

gpe_status gpe_step_observe(gpe_solver *s, uint64_t nsteps, double out[4]) {
    if (!s || !out || nsteps == 0) {
        set_error("gpe_step_observe: invalid argument");
        return GPE_EINVAL;
    }
    if (s->cfg.mode != GPE_REAL_TIME) {
        set_error("gpe_step_observe: real time only (imaginary time renormalises every step)");
        return GPE_ENOTSUP;
    }
    if (nsteps > 1) {
        const gpe_status st = s->step(nsteps - 1, nullptr);
        if (st != GPE_OK) return st;
    } else if (!s->odd_valid) {
        set_error("gpe_step_observe: real-time stepping needs an odd field (gpe_set_state or gpe_init_odd)");
        return GPE_EINVAL;
    }
    if (s->cfg.potential == GPE_POT_ARRAY && !s->pot) {
        set_error("gpe_step_observe: potential kind is ARRAY but no potential was set");
        return GPE_EINVAL;
    }
    // the last step as two gather passes (any family gives the same bits),
    // the odd one accumulating the observables of the even field it reads
    const int tile =
        s->kernel == GPE_KERNEL_GATHER ? s->tile : default_tile(GPE_KERNEL_GATHER, s->cfg.precision, MODE_REAL, s->L);
    // the observables are summed per cell: a node or class-group tile hands the
    // odd pass to the default one-thread-per-cell tile
    const int obs_tile =
        tile_shape(GPE_KERNEL_GATHER, tile).node ? cell_default_tile(s->cfg.precision, MODE_REAL, s->L) : tile;
    const gpe_status st = s->cfg.precision == GPE_FP64 ? s->observe_step<double, FORM_REF>(obs_tile, tile)
                                                       : s->observe_step<float, FORM_DIFF>(obs_tile, tile);
    if (st != GPE_OK) return st;
    GPE_HIP_TRY(hipMemcpyAsync(out, s->scal + 4, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    return GPE_OK;
}

gpe_status gpe_compare_fields(gpe_solver *a, int field_a, gpe_solver *b, int field_b, double out[3]) {
    if (!a || !b || !out || (field_a != 0 && field_a != 1) || (field_b != 0 && field_b != 1)) {
        set_error("gpe_compare_fields: invalid argument");
        return GPE_EINVAL;
    }
    if (a->cfg.nx != b->cfg.nx || a->cfg.ny != b->cfg.ny || a->cfg.nz != b->cfg.nz) {
        set_error("gpe_compare_fields: the solvers' grids differ");
        return GPE_EINVAL;
    }
    const int fa = field_a == 0 ? a->ie : a->io, fb = field_b == 0 ? b->ie : b->io;
    GPE_HIP_TRY(hipStreamSynchronize(b->stream));  // b's field is read on a's stream
    const bool a64 = a->cfg.precision == GPE_FP64, b64 = b->cfg.precision == GPE_FP64;
    if (a64 && b64) launch_distance<double, double>(a, fa, b, fb);
    else if (a64) launch_distance<double, float>(a, fa, b, fb);
    else if (b64) launch_distance<float, double>(a, fa, b, fb);
    else launch_distance<float, float>(a, fa, b, fb);
    GPE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_finalize<3>), dim3(1), dim3(256), 0, a->stream, a->partials, obs_blocks(a->L), 1.0, 0,
                       a->scal + 4);
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipMemcpyAsync(out, a->scal + 4, 3 * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    GPE_HIP_TRY(hipStreamSynchronize(a->stream));
    return GPE_OK;
}

gpe_status gpe_get_field_layout(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t precision, gpe_field_layout *out) {
    if (!out || nx == 0 || ny == 0 || nz == 0 || nx > 65536 || ny > 65536 || nz > 65536 || precision > GPE_FP32_BULK) {
        set_error("gpe_get_field_layout: invalid argument");
        return GPE_EINVAL;
    }
    const size_t eb = precision == GPE_FP64 ? 16 : 8;
    const Layout L = make_layout(nx, ny, nz, eb);
    out->elems = (uint64_t)L.elems;
    out->origin = (uint64_t)L.origin;
    out->xp = (uint64_t)L.xp;
    out->sy = (uint64_t)L.sy;
    out->sz = (uint64_t)L.sz;
    out->elem_bytes = (uint32_t)eb;
    out->reserved = 0;
    return GPE_OK;
}

gpe_status gpe_get_field_device(gpe_solver *s, int field, void **dev) {
    if (!s || !dev || (field != 0 && field != 1)) {
        set_error("gpe_get_field_device: invalid argument");
        return GPE_EINVAL;
    }
    *dev = s->field[field == 0 ? s->ie : s->io];
    return GPE_OK;
}

gpe_status gpe_observables_noalloc(const gpe_config *cfg, const void *field, const void *pot_dt, double *out,
                                   void *stream) {
    if (!cfg || !field || !out) {
        set_error("gpe_observables_noalloc: null argument");
        return GPE_EINVAL;
    }
    const gpe_config &c = *cfg;
    if (c.nx == 0 || c.ny == 0 || c.nz == 0 || c.nx > 65536 || c.ny > 65536 || c.nz > 65536 ||
        c.precision > GPE_FP32_BULK || c.potential > GPE_POT_TRAP || !(c.block_scale > 0) || !(c.dt > 0) ||
        !std::isfinite(c.g) || !std::isfinite(c.trap_ratio)) {
        set_error("gpe_observables_noalloc: invalid configuration");
        return GPE_EINVAL;
    }
    if (c.potential == GPE_POT_ARRAY && !pot_dt) {
        set_error("gpe_observables_noalloc: potential kind is ARRAY but pot_dt is NULL");
        return GPE_EINVAL;
    }
    const Layout L = make_layout(c.nx, c.ny, c.nz, c.precision == GPE_FP64 ? 16 : 8);
    if (obs_blocks(L) > kNoAllocBlocks) {
        set_error("gpe_observables_noalloc: %u x %u x %u cells need %lld blocks of partials, more than the %lld "
                  "the scratch holds",
                  c.nx, c.ny, c.nz, obs_blocks(L), kNoAllocBlocks);
        return GPE_EINVAL;
    }
    double *partials = nullptr;
    const gpe_status st = scratch_address(&partials);
    if (st != GPE_OK) return st;
    const hipStream_t hs = (hipStream_t)stream;
    if (c.precision == GPE_FP64) return observables_noalloc_t<double>(c, L, field, pot_dt, out, hs, partials);
    return observables_noalloc_t<float>(c, L, field, pot_dt, out, hs, partials);
}

gpe_status gpe_synchronize(gpe_solver *s) {
    if (!s) return GPE_EINVAL;
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    return GPE_OK;
}

gpe_status gpe_set_variant(gpe_solver *s, uint32_t kernel, uint32_t tile) {
    if (!s || kernel == GPE_KERNEL_AUTO || kernel > GPE_KERNEL_FUSED) {
        set_error("gpe_set_variant: invalid kernel family");
        return GPE_EINVAL;
    }
    if ((int)tile > num_tiles((int)kernel)) {
        set_error("gpe_set_variant: tile %u out of range", tile);
        return GPE_EINVAL;
    }
    const int t0 = tile == 0 ? default_tile((int)kernel, s->cfg.precision, MODE_REAL, s->L) : (int)tile - 1;
    if (!tile_supported((int)kernel, t0, s->cfg.precision, s->L)) {
        set_error("gpe_set_variant: fused tile %u does not fit LDS in FP64, or is a row tile on a field of 2 GiB "
                  "or more", tile);
        return GPE_EINVAL;
    }
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    if (kernel == GPE_KERNEL_FUSED) {
        for (int f = 0; f < 4; f++) {
            gpe_status st = s->alloc_field(f);
            if (st != GPE_OK) return st;
        }
    }
    s->kernel = (int)kernel;
    s->tile = t0;
    s->tile_default = tile == 0;
    return GPE_OK;
}

int gpe_variant_tiles(uint32_t kernel) {
    if (kernel >= GPE_KERNEL_GATHER && kernel <= GPE_KERNEL_FUSED) return num_tiles((int)kernel);
    return 0;
}

void gpe_variant_shape(uint32_t kernel, uint32_t tile, int32_t out[3]) {
    out[0] = out[1] = out[2] = 0;
    if (kernel >= GPE_KERNEL_GATHER && kernel <= GPE_KERNEL_FUSED && tile >= 1 && (int)tile <= num_tiles((int)kernel)) {
        const TileShape t = tile_shape((int)kernel, (int)tile - 1);
        out[0] = t.ty;
        out[1] = t.cz;
        out[2] = (t.remap == 1 ? 1 : 0) | (t.nt << 1) | (t.remap == 2 ? 4 : 0) | (t.row << 3) | (t.node << 4);  // bits 4-5
    }
}

gpe_status gpe_get_info(const gpe_solver *s, gpe_solver_info *info) {
    if (!s || !info) {
        set_error("gpe_get_info: null argument");
        return GPE_EINVAL;
    }
    info->nodes = s->nodes();
    info->device_bytes = s->device_bytes();
    info->steps_done = s->steps_done;
    int k = s->kernel, t = s->tile;
    if (s->cfg.mode == GPE_IMAG_TIME) s->imag_variant(k, t);
    info->kernel = (uint32_t)k;
    info->tile = (uint32_t)t + 1;
    info->launches_per_step = (s->cfg.mode == GPE_REAL_TIME && s->kernel != GPE_KERNEL_FUSED) ? 2 : 1;
    const double S = (double)s->elem_bytes();
    const double P = s->cfg.potential == GPE_POT_ARRAY ? S / 2 : 0.0;
    // read prev, read next, write next (+ streamed V) per node-update, two
    // node-updates per node-step.
    info->bytes_per_node_step = 2.0 * (3.0 * S + P);
    return GPE_OK;
}

}  // extern "C"
