// gpe_batch.hip -- many independent condensates stepped by shared launches
// (include/vaikkari_gpe_batch.h).
//
// Device side: every active instance has a descriptor (layout, trap,
// coefficients, tile grid, field pointers, steps left, ping-pong parity).
// Per step, shared kernels run over the work items (instance, tile) of all
// instances: the two in-place leapfrog half-updates (gather pass) over the
// real-time instances, the z-march Euler update over the imaginary-time
// ones, and the per-instance FP64 norm reduction of the latter.  Work items
// are laid out instance-major, one block each (the kernels loop when there
// are more than the grid cap); a table maps each work item to its instance,
// so a block finds its descriptor with one load.  An "advance" kernel at the
// end of a chunk updates steps-left and parity on the device, so a captured
// chunk graph can be replayed without host traffic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "gpe_internal.hpp"
#include "gpe_hip_check.hpp"
#include "gpe_kernels.hpp"
#include "gpe_solver.hpp"
#include "../../include/vaikkari_gpe_batch.h"

using namespace gpe;

namespace gpe {

template <typename T>
struct InstDesc {
    Layout L;
    Trap trap;
    Hamiltonian<T, sizeof(T) == 8 ? FORM_REF : FORM_DIFF> H;
    double dt;
    Grid3 gg;      // real time (gather tiles, in-place leapfrog)
    Grid3 zg;      // imaginary time (z-march tiles)
    int zcz;
    // device addresses (see global_ptr): the fields, the per-tile norm
    // partials and the scalars [0] N, [1] read scale
    unsigned long long buf[4];
    unsigned long long partials;
    unsigned long long scal;
    double w;
    int tile0;     // first work item in its kernel's list
    int ntiles;
    long long steps_left;
    int parity;
};

struct BatchHeader {
    int n_real, total_real, n_imag, total_imag;
};

// Blocks per batched launch: one per work item, up to this many (the kernels
// loop over the rest).
constexpr int kMaxBatchGrid = 1 << 20;

// A pointer read from memory is a generic pointer: the compiler cannot tell
// that it addresses global memory and accesses it with flat instructions.
// The descriptors therefore hold the fields' addresses as integers, made into
// global-address-space pointers here, so the accesses compile to global loads
// and stores as they do for kernel arguments.
template <typename P>
__device__ __forceinline__ P *global_ptr(unsigned long long a) {
    return (P *)(__attribute__((address_space(1))) P *)a;
}

// Tile of an instance's work item lt.  Every instance's items start at a
// multiple of 8 and work items go to the 8 XCDs round-robin, so item lt runs
// on XCD (lt mod 8) (or on XCD 7 - (lt mod 8) when the items are walked
// backwards).  With remap 2 (block rows divisible by 8) band (lt mod 8) is a
// y-band of every layer, as in block_coords: each XCD keeps the halo rows of
// its band in its own L2.  With remap 1 XCD (lt mod 8) takes the (lt mod 8)-th
// contiguous eighth of the tiles in (bx, by, bz) order, a slab of layers, so
// only the layers at the slab ends are shared between L2s.  Otherwise tiles go
// in (bx, by, bz) order.  Returns false for padding items (the item count is
// the tile count rounded up to a multiple of 8).
__device__ __forceinline__ bool item_tile(const Grid3 &G, int lt, int &bx, int &by, int &bz) {
    const int n = G.gx * G.gy * G.gz;
    int t = lt;
    if (G.remap == 2) {
        const int br = G.gy >> 3, bs = G.gx * br, slot = lt >> 3;
        const int layer = slot / bs, within = slot - layer * bs;
        bz = layer;
        by = (lt & 7) * br + within / G.gx;
        bx = within % G.gx;
        return layer < G.gz;
    }
    if (G.remap == 1) t = (lt & 7) * ((n + 7) >> 3) + (lt >> 3);
    bx = t % G.gx;
    by = (t / G.gx) % G.gy;
    bz = t / (G.gx * G.gy);
    return t < n;
}

constexpr int round8(int n) { return (n + 7) & ~7; }
// Field bytes per group of real-time instances.  A group's leapfrog reads and
// writes its fields twice per step; the second half-update walks the tiles
// backwards, so what the 256 MB Infinity Cache holds of the first is read
// again from there.  256 MiB keeps a group in that cache for the default
// cached tile (a streamed tile measured best at 512-640 MiB).
constexpr uint64_t kDefaultGroupBytes = 256ull << 20;
// Most kernel launches one captured chunk may hold (longer chunks are split).
constexpr uint64_t kMaxGraphLaunches = 16384;
constexpr int kImagTY = 4;
// Default tile of the real-time pass: one thread per cell and group of 4
// classes, 3 waves per block, cached read-modify-write, each XCD a contiguous
// eighth of the group (with 256 MiB groups the best measured on uniform
// batches of 16^3-48^3 cells, the 64-instance mixed form (+10 to +14 % over
// one thread per cell) and a serving stream's resident set (+3 %)).
constexpr int kRealTY = 3;
constexpr int kFinalizeThreads = 256;


// One half of a leapfrog step over all tiles of all real-time instances:
// half 0 updates odd from even, half 1 even from odd (in place, as
// GPUGrossPitaevskii_BCC.cu:478,480).
// The launch covers the work items [item0, item0 + nitems) of one group of
// instances (item0 and nitems are multiples of 8).
// GS (group slabs): block b runs on XCD b mod 8 (work-groups go to the XCDs
// round-robin), and XCD x takes the x-th contiguous eighth of the group's
// work items in order -- whole instances, layer after layer -- so the
// neighbour rows and layers a tile reads were fetched into that XCD's L2 by
// the tiles just before it; only the instances cut by an eighth's ends are
// shared between L2s.  The instances' own tile order is then the plain one.
// Half 1 walks each eighth backwards.  Otherwise half 1 walks all the work
// items backwards.
template <typename T, int TY, bool NT, bool GS>
__global__ void __launch_bounds__(64 * TY)
    k_batch_real(const InstDesc<T> *__restrict__ D, const int *__restrict__ items, int item0, int nitems, int s,
                 int chunk, int half) {
    constexpr int FORM = sizeof(T) == 8 ? FORM_REF : FORM_DIFF;
    for (int base = 0; base < nitems; base += gridDim.x) {
        // half 1 walks the work items backwards, starting where half 0 ended
        // (those fields are still in the caches)
        const int local = base + (int)blockIdx.x;
        if (local >= nitems) continue;
        int w;
        if (GS) {
            const int per = nitems >> 3, slot = local >> 3;
            w = item0 + (local & 7) * per + (half ? per - 1 - slot : slot);
        } else {
            w = item0 + (half ? nitems - 1 - local : local);
        }
        const int q = __builtin_amdgcn_readfirstlane(items[w]);
        const InstDesc<T> &d = D[q];
        const long long steps = d.steps_left < chunk ? d.steps_left : chunk;
        if (s >= steps) continue;
        int bx, by, bz;
        if (!item_tile(d.gg, w - d.tile0, bx, by, bz)) continue;  // padding to the next multiple of 8
        cplx_t<T> *next = global_ptr<cplx_t<T>>(d.buf[half == 0 ? 1 : 0]);
        const cplx_t<T> *prev = global_ptr<cplx_t<T>>(d.buf[half == 0 ? 0 : 1]);
        gather_tile<T, FORM, POT_TRAP, MODE_REAL, TY, NT>(next, prev, nullptr, d.L, d.H, d.trap, d.dt, nullptr, bx,
                                                         by, bz, d.gg.lsh);
    }
}

// The real-time half-update with one thread per cell and group of 4 classes
// (quad_update, as the single solver's class-group tiles): an instance's work
// item is a block of TY (group, cell rows) waves, groups fastest; d.gg holds
// the wave grid (gy = wave rows per layer) and d.ntiles the blocks.  Work
// items go to the XCDs as in k_batch_real; with d.gg.remap 1 XCD x takes the
// x-th contiguous eighth of an instance's blocks.
template <typename T, int TY, bool NT, bool GS>
__global__ void __launch_bounds__(64 * TY)
    k_batch_quad(const InstDesc<T> *__restrict__ D, const int *__restrict__ items, int item0, int nitems, int s,
                 int chunk, int half) {
    constexpr int FORM = sizeof(T) == 8 ? FORM_REF : FORM_DIFF;
    for (int base = 0; base < nitems; base += gridDim.x) {
        const int local = base + (int)blockIdx.x;
        if (local >= nitems) continue;
        int w;
        if (GS) {
            const int per = nitems >> 3, slot = local >> 3;
            w = item0 + (local & 7) * per + (half ? per - 1 - slot : slot);
        } else {
            w = item0 + (half ? nitems - 1 - local : local);
        }
        const int q = __builtin_amdgcn_readfirstlane(items[w]);
        const InstDesc<T> &d = D[q];
        const long long steps = d.steps_left < chunk ? d.steps_left : chunk;
        if (s >= steps) continue;
        const int lb = w - d.tile0, nb = d.ntiles;
        const int b = d.gg.remap == 1 ? (lb & 7) * ((nb + 7) >> 3) + (lb >> 3) : lb;
        if (b >= nb) continue;  // padding to the next multiple of 8
        const int wv = b * TY + (int)threadIdx.y;
        const int grp = __builtin_amdgcn_readfirstlane(wv % 3);
        int t = wv / 3;
        const int bx = t % d.gg.gx;
        t /= d.gg.gx;
        const int by = t % d.gg.gy, k = t / d.gg.gy;
        const int i = (bx << d.gg.lsh) + ((int)threadIdx.x & ((1 << d.gg.lsh) - 1));
        const int j = (by << (6 - d.gg.lsh)) + ((int)threadIdx.x >> d.gg.lsh);
        if (k >= d.L.nz || i >= d.L.nx || j >= d.L.ny) continue;
        cplx_t<T> *next = global_ptr<cplx_t<T>>(d.buf[half == 0 ? 1 : 0]);
        const cplx_t<T> *prev = global_ptr<cplx_t<T>>(d.buf[half == 0 ? 0 : 1]);
        if (grp == 0)
            quad_update<T, FORM, POT_TRAP, MODE_REAL, 0, NT>(next, prev, nullptr, d.L, d.H, d.trap, d.dt, (T)1, i, j, k);
        else if (grp == 1)
            quad_update<T, FORM, POT_TRAP, MODE_REAL, 1, NT>(next, prev, nullptr, d.L, d.H, d.trap, d.dt, (T)1, i, j, k);
        else
            quad_update<T, FORM, POT_TRAP, MODE_REAL, 2, NT>(next, prev, nullptr, d.L, d.H, d.trap, d.dt, (T)1, i, j, k);
    }
}

template <typename T>
__global__ void __launch_bounds__(64 * kImagTY)
    k_batch_imag(const InstDesc<T> *__restrict__ D, const int *__restrict__ items, const BatchHeader *__restrict__ hdr,
                 int s, int chunk) {
    constexpr int FORM = sizeof(T) == 8 ? FORM_REF : FORM_DIFF;
    __shared__ double lds[kImagTY];
    const int total = hdr->total_imag;
    for (int base = 0; base < total; base += gridDim.x) {
        // odd steps of a chunk walk the work items backwards (start where the last step ended)
        const int w = (s & 1) ? total - 1 - (base + (int)blockIdx.x) : base + (int)blockIdx.x;
        if (w < 0 || w >= total) continue;
        const int q = __builtin_amdgcn_readfirstlane(items[w]);
        const InstDesc<T> &d = D[q];
        const long long steps = d.steps_left < chunk ? d.steps_left : chunk;
        if (s >= steps) continue;
        const int p = (d.parity + s) & 1;
        int bx, by, bz;
        if (!item_tile(d.zg, w - d.tile0, bx, by, bz)) continue;  // padding to the next multiple of 8
        double part[1];
        cplx_t<T> *out = global_ptr<cplx_t<T>>(d.buf[1 - p]);
        const cplx_t<T> *in = global_ptr<cplx_t<T>>(d.buf[p]);
        const double *scale = global_ptr<double>(d.scal) + 1;
        // one layer per tile (the default) or a z-march of zcz layers
        if (d.zcz == 1)
            part[0] = gather_tile<T, FORM, POT_TRAP, MODE_IMAG, kImagTY, true>(out, in, nullptr, d.L, d.H, d.trap,
                                                                               d.dt, scale, bx, by, bz, d.zg.lsh);
        else
            part[0] = zmarch_tile<T, FORM, POT_TRAP, MODE_IMAG, kImagTY, true>(out, in, nullptr, d.L, d.H, d.trap,
                                                                               d.dt, scale, bx, by, bz, d.zcz,
                                                                               d.zg.lsh);
        block_sum<1, kImagTY>(part, lds);
        // partials in logical tile order, as the single solver stores them
        if (threadIdx.x == 0 && threadIdx.y == 0)
            global_ptr<double>(d.partials)[bx + d.zg.gx * (by + d.zg.gy * bz)] = part[0];
        __syncthreads();  // lds is reused by the next work item
    }
}

// Per imaginary-time instance: N = w * sum(partials), scale = 1/sqrt(N), in a
// fixed order (the same as k_finalize<1>).
template <typename T>
__global__ void __launch_bounds__(kFinalizeThreads)
    k_batch_finalize(const InstDesc<T> *__restrict__ D, const BatchHeader *__restrict__ hdr, int s, int chunk) {
    __shared__ double red[kFinalizeThreads];
    const int n = hdr->n_imag;
    const int t = threadIdx.x;
    for (int q = blockIdx.x; q < n; q += gridDim.x) {
        const InstDesc<T> &d = D[q];
        const long long steps = d.steps_left < chunk ? d.steps_left : chunk;
        if (s >= steps) continue;
        red[t] = strided_neumaier(global_ptr<double>(d.partials), 1, t, kFinalizeThreads, d.ntiles);
        __syncthreads();
        for (int h = kFinalizeThreads / 2; h > 0; h >>= 1) {
            if (t < h) red[t] += red[t + h];
            __syncthreads();
        }
        if (t == 0) {
            const double N = d.w * red[0];
            double *sc = global_ptr<double>(d.scal);
            sc[0] = N;
            sc[1] = 1.0 / sqrt(N);
        }
        __syncthreads();
    }
}

// End of a chunk: steps_left -= steps taken, parity ^= steps & 1.
template <typename T>
__global__ void k_batch_advance(InstDesc<T> *__restrict__ R, InstDesc<T> *__restrict__ I,
                                const BatchHeader *__restrict__ hdr, int chunk) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    InstDesc<T> *d = nullptr;
    if (q < hdr->n_real) d = &R[q];
    else if (q - hdr->n_real < hdr->n_imag) d = &I[q - hdr->n_real];
    if (!d) return;
    const long long steps = d->steps_left < chunk ? d->steps_left : chunk;
    d->steps_left -= steps;
    d->parity = (int)((d->parity + steps) & 1);
}

}  // namespace gpe

namespace {


struct BInst {
    gpe_instance_config cfg{};
    Layout L{};
    Trap trap{};
    Coefs co{};
    Grid3 gg{};
    Grid3 zg{};
    int zcz = 1;
    void *buf[4] = {nullptr, nullptr, nullptr, nullptr};
    double *partials = nullptr;
    double *scal = nullptr;
    int ntiles = 0;
    uint64_t steps_done = 0, steps_left = 0;
    int parity = 0;
    bool active = false;
    size_t bytes = 0;
};

}  // namespace

struct gpe_batch {
    uint32_t precision = GPE_FP32_BULK;
    uint64_t budget = 0;
    int real_ty = kRealTY;  // waves per block of the real-time pass (fixed while instances exist)
    bool real_nt = false;   // streaming (non-temporal) read-modify-write of the updated field
    bool real_quad = true;  // one thread per cell and group of 4 classes (real_ty 3, 6 or 12 waves)
    int real_order = 2;       // tile order over the XCDs: 0 y-bands (or plain), 1 a contiguous eighth
                              // of every instance's tiles, 2 of the whole group's work items
    // Real-time instances advance a whole chunk group by group: contiguous
    // runs of instances whose fields total at most group_bytes (0: one group)
    uint64_t group_bytes = kDefaultGroupBytes;
    std::vector<std::pair<int, int>> real_groups;  // (first work item, work items) per group
    std::vector<BInst> inst;
    uint64_t resident = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device descriptor arrays (capacity grows by doubling)
    void *d_real = nullptr, *d_imag = nullptr;
    BatchHeader *d_hdr = nullptr;
    int capacity = 0;
    void *h_stage = nullptr;  // pinned staging for descriptor uploads
    size_t stage_bytes = 0;
    int n_real = 0, n_imag = 0;
    int grid_imag = 1;                 // blocks of the imaginary-time launch
    int grid_real_items = 0;           // real-time work items (offset of the imaginary-time ones)
    // instance of every work item: real-time items [0, total_real), then imaginary-time ones
    int *d_items = nullptr;
    size_t items_cap = 0;
    std::vector<int> h_items;
    bool desc_dirty = true;
    // captured chunk graph
    hipGraphExec_t gexec = nullptr;
    uint32_t graph_chunk = 0;
    bool graph_dirty = true;
    bool use_graph = true;  // GPE_BATCH_GRAPH; false: GPE_BATCH_STREAM
    double *obs_partials = nullptr;
    long long obs_cap = 0;

    size_t elem() const { return precision == GPE_FP64 ? 16 : 8; }
    ~gpe_batch();
};

gpe_batch::~gpe_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto &b : inst) {
        for (auto *p : b.buf)
            if (p) (void)hipFree(p);
        if (b.partials) (void)hipFree(b.partials);
        if (b.scal) (void)hipFree(b.scal);
    }
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (d_real) (void)hipFree(d_real);
    if (d_imag) (void)hipFree(d_imag);
    if (d_hdr) (void)hipFree(d_hdr);
    if (d_items) (void)hipFree(d_items);
    if (h_stage) (void)hipHostFree(h_stage);
    if (obs_partials) (void)hipFree(obs_partials);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

bool valid_instance(const gpe_instance_config *c) {
    return c && c->nx > 0 && c->ny > 0 && c->nz > 0 && c->nx <= 4096 && c->ny <= 4096 && c->nz <= 4096 &&
           c->mode <= GPE_IMAG_TIME && c->block_scale > 0 && c->dt > 0 && std::isfinite(c->g) &&
           std::isfinite(c->trap_ratio) && c->init <= GPE_INIT_GAUSSIAN;
}

size_t instance_bytes(const gpe_instance_config *c, size_t elem, int real_ty, int real_order, bool real_quad,
                      Layout *Lout, Grid3 *gg, Grid3 *zg, int *zcz, int *ntiles) {
    const Layout L = make_layout(c->nx, c->ny, c->nz, elem);
    size_t bytes = 2 * (size_t)L.elems * elem;  // even and odd (real time) or ping-pong pair (imaginary time)
    int tiles = 0;
    if (c->mode == GPE_REAL_TIME && real_quad) {
        // k_batch_quad: the wave grid, 3 group waves per wave row of cells
        Grid3 G{};
        G.lsh = row_lanes_log2(L.nx);
        G.gx = (L.nx + (1 << G.lsh) - 1) >> G.lsh;
        G.gy = (L.ny + (64 >> G.lsh) - 1) / (64 >> G.lsh);
        G.gz = L.nz;
        G.remap = real_order == 1 ? 1 : 0;
        tiles = (int)((3LL * G.gx * G.gy * G.gz + real_ty - 1) / real_ty);
        if (gg) *gg = G;
    } else if (c->mode == GPE_REAL_TIME) {
        Grid3 G{};
        G.lsh = row_lanes_log2(L.nx);
        G.gx = (L.nx + (1 << G.lsh) - 1) >> G.lsh;
        G.gy = (L.ny + real_ty * (64 >> G.lsh) - 1) / (real_ty * (64 >> G.lsh));
        G.gz = L.nz;
        // group slabs: the kernel spreads the group over the XCDs, each
        // instance's tiles in plain (layer-major) order
        G.remap = real_order == 2 ? 0 : real_order == 1 ? 1 : G.gy % 8 == 0 ? 2 : 0;
        tiles = G.gx * G.gy * G.gz;
        if (gg) *gg = G;
    } else {
        const int cz = c->zlayers ? (int)c->zlayers : 1;
        Grid3 G{};
        G.lsh = row_lanes_log2(L.nx);
        G.gx = (L.nx + (1 << G.lsh) - 1) >> G.lsh;
        G.gy = (L.ny + kImagTY * (64 >> G.lsh) - 1) / (kImagTY * (64 >> G.lsh));
        G.gz = (L.nz + cz - 1) / cz;
        G.remap = G.gy % 8 == 0 ? 2 : 0;
        tiles = G.gx * G.gy * G.gz;
        bytes += (size_t)tiles * sizeof(double) + 8 * sizeof(double);
        if (zg) *zg = G;
        if (zcz) *zcz = cz;
    }
    if (Lout) *Lout = L;
    if (ntiles) *ntiles = tiles;
    return bytes;
}

template <typename T>
void fill_desc(InstDesc<T> &d, const BInst &b) {
    std::memset(&d, 0, sizeof(d));
    d.L = b.L;
    d.trap = b.trap;
    d.H.lapfac = (T)b.co.lapfac_dt;
    d.H.lapfac0 = (T)b.co.lapfac0_dt;
    d.H.g = (T)b.co.g_dt;
    d.dt = b.co.dt;
    d.gg = b.gg;
    d.zg = b.zg;
    d.zcz = b.zcz;
    for (int f = 0; f < 4; f++) d.buf[f] = (unsigned long long)(uintptr_t)b.buf[f];
    d.partials = (unsigned long long)(uintptr_t)b.partials;
    d.scal = (unsigned long long)(uintptr_t)b.scal;
    d.w = b.co.w;
    d.ntiles = b.ntiles;
    d.steps_left = (long long)b.steps_left;
    d.parity = b.parity;
}

template <typename T>
gpe_status upload_desc(gpe_batch *B) {
    const size_t dsz = sizeof(InstDesc<T>);
    int nr = 0, ni = 0;
    for (auto &b : B->inst)
        if (b.active && b.steps_left > 0) (b.cfg.mode == GPE_REAL_TIME ? nr : ni)++;
    const int need = std::max(std::max(nr, ni), 1);
    if (need > B->capacity) {
        int cap = std::max(64, B->capacity);
        while (cap < need) cap *= 2;
        if (B->d_real) (void)hipFree(B->d_real);
        if (B->d_imag) (void)hipFree(B->d_imag);
        if (B->h_stage) (void)hipHostFree(B->h_stage);
        B->d_real = B->d_imag = B->h_stage = nullptr;
        B->capacity = 0;
        GPE_HIP_TRY(hipMalloc(&B->d_real, cap * dsz));
        GPE_HIP_TRY(hipMalloc(&B->d_imag, cap * dsz));
        B->stage_bytes = 2 * cap * dsz + sizeof(BatchHeader);
        GPE_HIP_TRY(hipHostMalloc(&B->h_stage, B->stage_bytes));
        B->capacity = cap;
        B->graph_dirty = true;  // descriptor addresses changed
    }
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));  // the staging buffer is reused
    InstDesc<T> *R = (InstDesc<T> *)B->h_stage;
    InstDesc<T> *I = R + B->capacity;
    BatchHeader *H = (BatchHeader *)(I + B->capacity);
    int tr = 0, ti = 0;
    nr = ni = 0;
    B->real_groups.clear();
    uint64_t group_fill = 0;
    for (auto &b : B->inst) {
        if (!b.active || b.steps_left == 0) continue;
        if (b.cfg.mode == GPE_REAL_TIME) {
            if (nr == 0 || (B->group_bytes && group_fill + b.bytes > B->group_bytes)) {
                B->real_groups.emplace_back(tr, 0);
                group_fill = 0;
            }
            group_fill += b.bytes;
            fill_desc(R[nr], b);
            R[nr].tile0 = tr;
            tr += round8(b.ntiles);
            B->real_groups.back().second = tr - B->real_groups.back().first;
            nr++;
        } else {
            fill_desc(I[ni], b);
            I[ni].tile0 = ti;
            ti += round8(b.ntiles);
            ni++;
        }
    }
    H->n_real = nr;
    H->total_real = tr;
    H->n_imag = ni;
    H->total_imag = ti;
    B->h_items.assign((size_t)std::max(tr + ti, 1), 0);
    for (int q = 0; q < nr; q++)
        std::fill(B->h_items.begin() + R[q].tile0, B->h_items.begin() + R[q].tile0 + round8(R[q].ntiles), q);
    for (int q = 0; q < ni; q++)
        std::fill(B->h_items.begin() + tr + I[q].tile0, B->h_items.begin() + tr + I[q].tile0 + round8(I[q].ntiles), q);
    if (B->h_items.size() > B->items_cap) {
        if (B->d_items) GPE_HIP_TRY(hipFree(B->d_items));
        B->d_items = nullptr;
        B->items_cap = 0;
        const size_t cap = std::max<size_t>(B->h_items.size() * 2, 4096);
        GPE_HIP_TRY(hipMalloc(&B->d_items, cap * sizeof(int)));
        B->items_cap = cap;
    }
    GPE_HIP_TRY(hipMemcpyAsync(B->d_items, B->h_items.data(), B->h_items.size() * sizeof(int),
                               hipMemcpyHostToDevice, B->stream));
    GPE_HIP_TRY(hipMemcpyAsync(B->d_real, R, (size_t)std::max(nr, 1) * dsz, hipMemcpyHostToDevice, B->stream));
    GPE_HIP_TRY(hipMemcpyAsync(B->d_imag, I, (size_t)std::max(ni, 1) * dsz, hipMemcpyHostToDevice, B->stream));
    GPE_HIP_TRY(hipMemcpyAsync(B->d_hdr, H, sizeof(BatchHeader), hipMemcpyHostToDevice, B->stream));
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    B->n_real = nr;
    B->n_imag = ni;
    B->grid_real_items = tr;
    B->grid_imag = std::max(1, std::min(ti, kMaxBatchGrid));
    B->desc_dirty = false;
    return GPE_OK;
}

template <typename T>
void enqueue_chunk(gpe_batch *B, uint32_t chunk, hipStream_t st) {
    const InstDesc<T> *R = (const InstDesc<T> *)B->d_real;
    const InstDesc<T> *I = (const InstDesc<T> *)B->d_imag;
    // Real-time instances: each group runs the whole chunk while its fields
    // are in the caches, then the next group.  Instances are independent, so
    // the order does not change any result.
    for (const auto &grp : B->real_groups) {
        const int grid = std::max(1, std::min(grp.second, kMaxBatchGrid));
        for (uint32_t s = 0; s < chunk; s++) {
            for (int half = 0; half < 2; half++) {
#define GPE_BATCH_REAL(TY, NT, GS)                                                                      \
    hipLaunchKernelGGL((k_batch_real<T, TY, NT, GS>), dim3(grid), dim3(64, TY), 0, st, R, B->d_items,     \
                       grp.first, grp.second, (int)s, (int)chunk, half)
                const bool gs = B->real_order == 2;
#define GPE_BATCH_QUAD(TY, NT, GS)                                                                      \
    hipLaunchKernelGGL((k_batch_quad<T, TY, NT, GS>), dim3(grid), dim3(64, TY), 0, st, R, B->d_items,     \
                       grp.first, grp.second, (int)s, (int)chunk, half)
                if (B->real_quad) {
                    switch (B->real_ty * 4 + (B->real_nt ? 2 : 0) + (gs ? 1 : 0)) {
                        case 12: GPE_BATCH_QUAD(3, false, false); break;
                        case 13: GPE_BATCH_QUAD(3, false, true); break;
                        case 14: GPE_BATCH_QUAD(3, true, false); break;
                        case 15: GPE_BATCH_QUAD(3, true, true); break;
                        case 24: GPE_BATCH_QUAD(6, false, false); break;
                        case 25: GPE_BATCH_QUAD(6, false, true); break;
                        case 26: GPE_BATCH_QUAD(6, true, false); break;
                        case 27: GPE_BATCH_QUAD(6, true, true); break;
                        case 48: GPE_BATCH_QUAD(12, false, false); break;
                        case 49: GPE_BATCH_QUAD(12, false, true); break;
                        case 50: GPE_BATCH_QUAD(12, true, false); break;
                        default: GPE_BATCH_QUAD(12, true, true);
                    }
                    continue;
                }
#undef GPE_BATCH_QUAD
                switch (B->real_ty * 4 + (B->real_nt ? 2 : 0) + (gs ? 1 : 0)) {
                    case 4: GPE_BATCH_REAL(1, false, false); break;
                    case 5: GPE_BATCH_REAL(1, false, true); break;
                    case 6: GPE_BATCH_REAL(1, true, false); break;
                    case 7: GPE_BATCH_REAL(1, true, true); break;
                    case 8: GPE_BATCH_REAL(2, false, false); break;
                    case 9: GPE_BATCH_REAL(2, false, true); break;
                    case 10: GPE_BATCH_REAL(2, true, false); break;
                    case 11: GPE_BATCH_REAL(2, true, true); break;
                    case 16: GPE_BATCH_REAL(4, false, false); break;
                    case 17: GPE_BATCH_REAL(4, false, true); break;
                    case 19: GPE_BATCH_REAL(4, true, true); break;
                    case 32: GPE_BATCH_REAL(8, false, false); break;
                    case 33: GPE_BATCH_REAL(8, false, true); break;
                    case 34: GPE_BATCH_REAL(8, true, false); break;
                    case 35: GPE_BATCH_REAL(8, true, true); break;
                    default: GPE_BATCH_REAL(4, true, false);
                }
#undef GPE_BATCH_REAL
            }
        }
    }
    for (uint32_t s = 0; B->n_imag > 0 && s < chunk; s++) {
        hipLaunchKernelGGL((k_batch_imag<T>), dim3(B->grid_imag), dim3(64, kImagTY), 0, st, I,
                           B->d_items + B->grid_real_items, B->d_hdr, (int)s, (int)chunk);
        hipLaunchKernelGGL((k_batch_finalize<T>), dim3(std::min(B->n_imag, 1024)), dim3(kFinalizeThreads), 0, st, I,
                           B->d_hdr, (int)s, (int)chunk);
    }
    const int n = B->n_real + B->n_imag;
    hipLaunchKernelGGL((k_batch_advance<T>), dim3((n + 255) / 256), dim3(256), 0, st, (InstDesc<T> *)B->d_real,
                       (InstDesc<T> *)B->d_imag, B->d_hdr, (int)chunk);
}

// Kernel launches per step of a chunk under the grouping the next upload
// forms (the same walk as upload_desc).
uint64_t launches_per_step(const gpe_batch *B) {
    uint64_t groups = 0, fill = 0;
    bool imag = false;
    for (const auto &b : B->inst) {
        if (!b.active || b.steps_left == 0) continue;
        if (b.cfg.mode != GPE_REAL_TIME) {
            imag = true;
            continue;
        }
        if (groups == 0 || (B->group_bytes && fill + b.bytes > B->group_bytes)) {
            groups++;
            fill = 0;
        }
        fill += b.bytes;
    }
    return 2 * groups + (imag ? 2 : 0);
}

template <typename T>
gpe_status run_t(gpe_batch *B, uint32_t chunk, uint32_t chunks, double *elapsed_ms) {
    if (B->desc_dirty) {
        gpe_status st = upload_desc<T>(B);
        if (st != GPE_OK) return st;
        B->graph_dirty = true;  // instance counts are baked into the launch sequence
    }
    if (B->n_real + B->n_imag == 0) {
        if (elapsed_ms) *elapsed_ms = 0.0;
        return GPE_OK;
    }
    if (!B->use_graph) {
        if (elapsed_ms) GPE_HIP_TRY(hipEventRecord(B->ev0, B->stream));
        for (uint32_t c = 0; c < chunks; c++) enqueue_chunk<T>(B, chunk, B->stream);
        GPE_HIP_TRY(hipGetLastError());
        if (elapsed_ms) GPE_HIP_TRY(hipEventRecord(B->ev1, B->stream));
    } else if (B->graph_dirty || B->graph_chunk != chunk || !B->gexec) {
        if (B->gexec) {
            (void)hipGraphExecDestroy(B->gexec);
            B->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        GPE_HIP_TRY(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
        enqueue_chunk<T>(B, chunk, B->stream);
        GPE_HIP_TRY(hipStreamEndCapture(B->stream, &g));
        const hipError_t ie = hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            B->gexec = nullptr;
            set_error("gpe_batch_run: graph instantiation failed: %s", hipGetErrorString(ie));
            return GPE_EDEVICE;
        }
        B->graph_chunk = chunk;
        B->graph_dirty = false;
    }
    if (B->use_graph) {
        if (elapsed_ms) GPE_HIP_TRY(hipEventRecord(B->ev0, B->stream));
        for (uint32_t c = 0; c < chunks; c++) GPE_HIP_TRY(hipGraphLaunch(B->gexec, B->stream));
        if (elapsed_ms) GPE_HIP_TRY(hipEventRecord(B->ev1, B->stream));
    }
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    if (elapsed_ms) {
        float ms = 0.0f;
        GPE_HIP_TRY(hipEventElapsedTime(&ms, B->ev0, B->ev1));
        *elapsed_ms = ms;
    }
    // host mirror of k_batch_advance
    bool finished = false;
    for (auto &b : B->inst) {
        if (!b.active || b.steps_left == 0) continue;
        const uint64_t take = std::min<uint64_t>(b.steps_left, (uint64_t)chunk * chunks);
        b.steps_left -= take;
        b.steps_done += take;
        b.parity = (int)((b.parity + take) & 1);
        if (b.steps_left == 0) finished = true;
    }
    if (finished) B->desc_dirty = true;  // drop finished instances from the work lists
    return GPE_OK;
}

BInst *get_inst(gpe_batch *B, int32_t id) {
    if (!B || id < 0 || id >= (int32_t)B->inst.size() || !B->inst[id].active) {
        set_error("batch: no active instance %d", id);
        return nullptr;
    }
    return &B->inst[id];
}

// Scales the even field of a device-initialised instance so that its
// discrete norm w * sum |psi|^2 (FP64, on the device) is 1: the continuum
// amplitude (pi sigma^2)^(-3/4) of the recipe gives less than 1 on a box that
// clips the Gaussian.
gpe_status normalize_even(gpe_batch *B, int32_t id) {
    double obs[4];
    gpe_status st = gpe_batch_observables(B, id, obs);
    if (st != GPE_OK) return st;
    BInst &b = B->inst[id];
    if (!(obs[0] > 0) || !std::isfinite(obs[0])) {
        set_error("gpe_batch_add: initial state has no positive finite norm");
        return GPE_EUNSTABLE;
    }
    const double s = 1.0 / std::sqrt(obs[0]);
    if (B->precision == GPE_FP64)
        hipLaunchKernelGGL((k_scale_const<double>), dim3(1024), dim3(256), 0, B->stream, (double2 *)b.buf[0],
                           b.L.elems, s);
    else
        hipLaunchKernelGGL((k_scale_const<float>), dim3(1024), dim3(256), 0, B->stream, (float2 *)b.buf[0],
                           b.L.elems, s);
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    return GPE_OK;
}

// odd = even exp(i mu dt/2), the phase of the reference's start of the
// leapfrog (GPUGrossPitaevskii_BCC.cu:325,342-343).  mu = NaN: the state's own
// mu = (K + P + I)/N of the instance's even field (FP64 observables).
gpe_status start_odd(gpe_batch *B, int32_t id, double mu) {
    if (std::isnan(mu)) {
        double obs[4];
        gpe_status st = gpe_batch_observables(B, id, obs);
        if (st != GPE_OK) return st;
        mu = (obs[1] + obs[2] + obs[3]) / obs[0];
        if (!std::isfinite(mu)) {
            set_error("gpe_batch_add: initial state has no finite chemical potential");
            return GPE_EUNSTABLE;
        }
    }
    BInst &b = B->inst[id];
    const double ph = 0.5 * mu * b.cfg.dt;
    if (B->precision == GPE_FP64)
        hipLaunchKernelGGL((k_phase_copy<double>), dim3(1024), dim3(256), 0, B->stream, (double2 *)b.buf[1],
                           (const double2 *)b.buf[0], b.L.elems, std::cos(ph), std::sin(ph));
    else
        hipLaunchKernelGGL((k_phase_copy<float>), dim3(1024), dim3(256), 0, B->stream, (float2 *)b.buf[1],
                           (const float2 *)b.buf[0], b.L.elems, std::cos(ph), std::sin(ph));
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    return GPE_OK;
}

// The even field of the Gaussian recipe; normalize_even and start_odd follow.
template <typename T>
gpe_status init_gaussian(gpe_batch *B, BInst &b) {
    const double s = b.cfg.sigma > 0 ? b.cfg.sigma : 1.0;
    const double amp = std::pow(3.14159265358979323846 * s * s, -0.75);
    hipLaunchKernelGGL((k_init_gaussian<T>), dim3(1024), dim3(256), 0, B->stream, (cplx_t<T> *)b.buf[0],
                       (cplx_t<T> *)nullptr, b.L, b.trap, 1.0 / (2.0 * s * s), amp, b.cfg.noise,
                       (unsigned long long)b.cfg.seed, 1.0, 0.0);
    GPE_HIP_TRY(hipGetLastError());
    return GPE_OK;
}

}  // namespace

extern "C" {

gpe_status gpe_batch_create(uint32_t precision, uint64_t budget, gpe_batch **out) {
    if (!out || precision > GPE_FP32_BULK) {
        set_error("gpe_batch_create: invalid argument");
        return GPE_EINVAL;
    }
    *out = nullptr;
    auto *B = new (std::nothrow) gpe_batch();
    if (!B) return GPE_ENOMEM;
    B->precision = precision;
    B->budget = budget;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&B->ev0) != hipSuccess || hipEventCreate(&B->ev1) != hipSuccess ||
        hipMalloc(&B->d_hdr, sizeof(BatchHeader)) != hipSuccess) {
        set_error("gpe_batch_create: HIP setup failed");
        delete B;
        return GPE_EDEVICE;
    }
    *out = B;
    return GPE_OK;
}

void gpe_batch_destroy(gpe_batch *b) { delete b; }

uint64_t gpe_batch_instance_bytes(const gpe_batch *B, const gpe_instance_config *c) {
    if (!B || !valid_instance(c)) return 0;
    return instance_bytes(c, B->elem(), B->real_ty, B->real_order, B->real_quad, nullptr, nullptr, nullptr, nullptr,
                          nullptr);
}

gpe_status gpe_batch_add(gpe_batch *B, const gpe_instance_config *c, int32_t *id) {
    if (!B || !id || !valid_instance(c)) {
        set_error("gpe_batch_add: invalid instance configuration");
        return GPE_EINVAL;
    }
    BInst b;
    b.cfg = *c;
    b.bytes = instance_bytes(c, B->elem(), B->real_ty, B->real_order, B->real_quad, &b.L, &b.gg, &b.zg, &b.zcz,
                             &b.ntiles);
    if (B->budget && B->resident + b.bytes > B->budget) {
        set_error("gpe_batch_add: %zu bytes would exceed the budget (%llu of %llu resident)", b.bytes,
                  (unsigned long long)B->resident, (unsigned long long)B->budget);
        return GPE_EBUSY;
    }
    double origin[3];
    for (int a = 0; a < 3; a++) origin[a] = c->origin[a];
    if (!std::isfinite(origin[0]) || !std::isfinite(origin[1]) || !std::isfinite(origin[2])) {
        const uint32_t dims[3] = {c->nx, c->ny, c->nz};
        const double a = std::sqrt(8.0) * c->block_scale;
        for (int d = 0; d < 3; d++) origin[d] = -0.5 * a * dims[d];
    }
    {
        const uint32_t dims[3] = {c->nx, c->ny, c->nz};
        const double vmax = trap_vmax(dims, origin, c->block_scale, c->trap_ratio);
        if (!linear_step_stable(c->dt, c->block_scale, vmax)) {
            set_error("gpe_batch_add: dt = %g violates dt*(6/b^2 + max V) < 2 (max V = %g)", c->dt, vmax);
            return GPE_EUNSTABLE;
        }
    }
    b.trap = make_trap(origin, c->block_scale, c->trap_ratio);
    const double bs = c->block_scale;
    const double lapfac = -0.5 * kHodge / (bs * bs);
    b.co.lapfac_dt = lapfac * c->dt;
    b.co.lapfac0_dt = (double)GPE_FACES * (-lapfac) * c->dt;
    b.co.g_dt = c->g * c->dt;
    b.co.dt = c->dt;
    b.co.w = node_measure(bs);
    b.co.lap = kHodge / (bs * bs);
    b.co.g = c->g;
    const int nfields = 2;
    const size_t fbytes = (size_t)b.L.elems * B->elem();
    auto fail = [&](gpe_status st) {
        (void)hipStreamSynchronize(B->stream);  // nothing enqueued above may still use these buffers
        (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
        for (auto *&p : b.buf)
            if (p) (void)hipFree(p);
        if (b.partials) (void)hipFree(b.partials);
        if (b.scal) (void)hipFree(b.scal);
        return st;
    };
    for (int f = 0; f < nfields; f++) {
        if (hipMalloc(&b.buf[f], fbytes) != hipSuccess) {
            b.buf[f] = nullptr;
            set_error("gpe_batch_add: out of device memory");
            return fail(GPE_ENOMEM);
        }
        if (hipMemsetAsync(b.buf[f], 0, fbytes, B->stream) != hipSuccess) return fail(GPE_EDEVICE);
    }
    // scal's initial value: lives on this frame until the stream is synchronised below
    const double scal_init[8] = {1.0, 1.0, 0, 0, 0, 0, 0, 0};
    if (c->mode == GPE_IMAG_TIME) {
        if (hipMalloc(&b.partials, (size_t)b.ntiles * sizeof(double)) != hipSuccess ||
            hipMalloc(&b.scal, 8 * sizeof(double)) != hipSuccess) {
            set_error("gpe_batch_add: out of device memory");
            return fail(GPE_ENOMEM);
        }
        if (hipMemcpyAsync(b.scal, scal_init, sizeof(scal_init), hipMemcpyHostToDevice, B->stream) != hipSuccess)
            return fail(GPE_EDEVICE);
    }
    b.steps_left = c->steps;
    b.active = true;
    if (c->init == GPE_INIT_GAUSSIAN) {
        const gpe_status st = B->precision == GPE_FP64 ? init_gaussian<double>(B, b) : init_gaussian<float>(B, b);
        if (st != GPE_OK) return fail(st);
    }
    if (hipStreamSynchronize(B->stream) != hipSuccess) {
        set_error("gpe_batch_add: initialisation failed");
        return fail(GPE_EDEVICE);
    }
    // reuse a free slot so ids stay small
    int32_t slot = -1;
    for (size_t q = 0; q < B->inst.size(); q++)
        if (!B->inst[q].active && !B->inst[q].buf[0]) {
            slot = (int32_t)q;
            break;
        }
    if (slot < 0) {
        B->inst.push_back(b);
        slot = (int32_t)B->inst.size() - 1;
    } else {
        B->inst[slot] = b;
    }
    B->resident += b.bytes;
    B->desc_dirty = true;
    *id = slot;
    if (c->init == GPE_INIT_GAUSSIAN) {
        // unit norm on the grid, then the odd level (real time; mu0 = NaN: the
        // state's own chemical potential)
        gpe_status st = normalize_even(B, slot);
        if (st == GPE_OK && c->mode == GPE_REAL_TIME) st = start_odd(B, slot, c->mu0);
        if (st != GPE_OK) {
            (void)gpe_batch_remove(B, slot);
            *id = -1;
            return st;
        }
    }
    return GPE_OK;
}

gpe_status gpe_batch_remove(gpe_batch *B, int32_t id) {
    BInst *b = get_inst(B, id);
    if (!b) return GPE_EINVAL;
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    for (auto *&p : b->buf)
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
    if (b->partials) (void)hipFree(b->partials);
    if (b->scal) (void)hipFree(b->scal);
    b->partials = nullptr;
    b->scal = nullptr;
    b->active = false;
    B->resident -= b->bytes;
    B->desc_dirty = true;
    return GPE_OK;
}

gpe_status gpe_batch_set_state(gpe_batch *B, int32_t id, const double *even, const double *odd) {
    BInst *b = get_inst(B, id);
    if (!b || !even || (b->cfg.mode == GPE_REAL_TIME && !odd)) {
        set_error("gpe_batch_set_state: invalid argument");
        return GPE_EINVAL;
    }
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    const int pe = b->cfg.mode == GPE_REAL_TIME ? 0 : b->parity;
    const bool f64 = B->precision == GPE_FP64;
    gpe_status st = f64 ? upload_field<double>(b->buf[pe], even, b->L, B->stream)
                        : upload_field<float>(b->buf[pe], even, b->L, B->stream);
    if (st == GPE_OK && b->cfg.mode == GPE_REAL_TIME)
        st = f64 ? upload_field<double>(b->buf[pe + 1], odd, b->L, B->stream)
                 : upload_field<float>(b->buf[pe + 1], odd, b->L, B->stream);
    if (b->scal) {
        const double init[2] = {1.0, 1.0};
        GPE_HIP_TRY(hipMemcpy(b->scal, init, sizeof(init), hipMemcpyHostToDevice));
    }
    return st;
}

gpe_status gpe_batch_get_state(gpe_batch *B, int32_t id, double *even, double *odd) {
    BInst *b = get_inst(B, id);
    if (!b) return GPE_EINVAL;
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    const int pe = b->cfg.mode == GPE_REAL_TIME ? 0 : b->parity;
    if (b->cfg.mode == GPE_IMAG_TIME && b->steps_done > 0) {
        // apply the pending renormalisation, as gpe_step does at its end
        const long long n = b->L.elems;
        if (B->precision == GPE_FP64)
            hipLaunchKernelGGL((k_scale_field<double>), dim3(1024), dim3(256), 0, B->stream, (double2 *)b->buf[pe], n,
                               b->scal + 1);
        else
            hipLaunchKernelGGL((k_scale_field<float>), dim3(1024), dim3(256), 0, B->stream, (float2 *)b->buf[pe], n,
                               b->scal + 1);
        GPE_HIP_TRY(hipGetLastError());
        const double one = 1.0;
        GPE_HIP_TRY(hipMemcpyAsync(b->scal + 1, &one, sizeof(double), hipMemcpyHostToDevice, B->stream));
        GPE_HIP_TRY(hipStreamSynchronize(B->stream));
        B->desc_dirty = true;
    }
    const bool f64 = B->precision == GPE_FP64;
    gpe_status st = GPE_OK;
    if (even)
        st = f64 ? download_field<double>(b->buf[pe], even, b->L, B->stream)
                 : download_field<float>(b->buf[pe], even, b->L, B->stream);
    if (st == GPE_OK && odd && b->cfg.mode == GPE_REAL_TIME)
        st = f64 ? download_field<double>(b->buf[pe + 1], odd, b->L, B->stream)
                 : download_field<float>(b->buf[pe + 1], odd, b->L, B->stream);
    return st;
}

gpe_status gpe_batch_observables(gpe_batch *B, int32_t id, double out[4]) {
    BInst *b = get_inst(B, id);
    if (!b || !out) return GPE_EINVAL;
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    const long long nb = (long long)((b->L.nx + 63) / 64) * ((b->L.ny + 3) / 4) * b->L.nz;
    if (nb > B->obs_cap) {
        if (B->obs_partials) (void)hipFree(B->obs_partials);
        B->obs_partials = nullptr;
        GPE_HIP_TRY(hipMalloc(&B->obs_partials, (size_t)nb * 4 * sizeof(double) + 8 * sizeof(double)));
        B->obs_cap = nb;
    }
    const int pe = b->cfg.mode == GPE_REAL_TIME ? 0 : b->parity;
    const dim3 grid((b->L.nx + 63) / 64, (b->L.ny + 3) / 4, b->L.nz), block(64, 4);
    double *res = B->obs_partials + nb * 4;
    if (B->precision == GPE_FP64)
        hipLaunchKernelGGL((k_observables<double, POT_TRAP, 4>), grid, block, 0, B->stream,
                           (const double2 *)b->buf[pe], nullptr, b->L, b->trap, 1.0 / b->co.dt, b->co.lap, b->co.g,
                           B->obs_partials);
    else
        hipLaunchKernelGGL((k_observables<float, POT_TRAP, 4>), grid, block, 0, B->stream,
                           (const float2 *)b->buf[pe], nullptr, b->L, b->trap, 1.0 / b->co.dt, b->co.lap, b->co.g,
                           B->obs_partials);
    GPE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_finalize<4>), dim3(1), dim3(256), 0, B->stream, B->obs_partials, nb, b->co.w, 0, res);
    GPE_HIP_TRY(hipGetLastError());
    double tmp[4];
    GPE_HIP_TRY(hipMemcpyAsync(tmp, res, sizeof(tmp), hipMemcpyDeviceToHost, B->stream));
    GPE_HIP_TRY(hipStreamSynchronize(B->stream));
    double sc = 1.0;
    if (b->scal && b->steps_done > 0) {
        // pending renormalisation of an imaginary-time field: N, K, P scale
        // with s^2, I with s^4
        GPE_HIP_TRY(hipMemcpy(&sc, b->scal + 1, sizeof(double), hipMemcpyDeviceToHost));
    }
    const double s2 = sc * sc;
    out[0] = tmp[0] * s2;
    out[1] = tmp[1] * s2;
    out[2] = tmp[2] * s2;
    out[3] = tmp[3] * s2 * s2;
    return GPE_OK;
}

gpe_status gpe_batch_instance(const gpe_batch *B, int32_t id, gpe_instance_info *info) {
    if (!B || !info || id < 0 || id >= (int32_t)B->inst.size()) {
        set_error("gpe_batch_instance: invalid argument");
        return GPE_EINVAL;
    }
    const BInst &b = B->inst[id];
    info->steps_done = b.steps_done;
    info->steps_left = b.steps_left;
    info->nodes = (uint64_t)b.cfg.nx * b.cfg.ny * b.cfg.nz * GPE_CLASSES;
    info->device_bytes = b.bytes;
    info->mode = b.cfg.mode;
    info->active = b.active ? 1u : 0u;
    return GPE_OK;
}

gpe_status gpe_batch_run(gpe_batch *B, uint32_t chunk_steps, uint32_t chunks, double *elapsed_ms) {
    if (!B || chunk_steps == 0) {
        set_error("gpe_batch_run: invalid argument");
        return GPE_EINVAL;
    }
    auto run = [B](uint32_t c, uint32_t n, double *ms) {
        return B->precision == GPE_FP64 ? run_t<double>(B, c, n, ms) : run_t<float>(B, c, n, ms);
    };
    // A captured chunk holds, for every step, one launch per group and half-step
    // plus two for the imaginary-time instances.  A long chunk over many groups
    // runs as shorter chunks of at most kMaxGraphLaunches launches each: where
    // chunks end changes no result (every instance still takes the same steps).
    const uint64_t per_step = launches_per_step(B);
    if (!B->use_graph || per_step * chunk_steps <= kMaxGraphLaunches)
        return run(chunk_steps, chunks, elapsed_ms);
    const uint32_t piece = (uint32_t)std::max<uint64_t>(1, kMaxGraphLaunches / per_step);
    const uint64_t total = (uint64_t)chunk_steps * chunks;
    uint64_t full = total / piece;
    const uint32_t rest = (uint32_t)(total % piece);
    double sum = 0.0, ms = 0.0;
    while (full > 0) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(full, 0xffffffffu);
        const gpe_status st = run(piece, n, elapsed_ms ? &ms : nullptr);
        if (st != GPE_OK) return st;
        sum += ms;
        full -= n;
    }
    if (rest) {
        const gpe_status st = run(rest, 1, elapsed_ms ? &ms : nullptr);
        if (st != GPE_OK) return st;
        sum += ms;
    }
    if (elapsed_ms) *elapsed_ms = sum;
    return GPE_OK;
}

uint64_t gpe_batch_chunk_node_steps(const gpe_batch *B, uint32_t chunk_steps) {
    if (!B) return 0;
    uint64_t t = 0;
    for (auto &b : B->inst)
        if (b.active)
            t += (uint64_t)b.cfg.nx * b.cfg.ny * b.cfg.nz * GPE_CLASSES *
                 std::min<uint64_t>(b.steps_left, chunk_steps);
    return t;
}

int32_t gpe_batch_active(const gpe_batch *B) {
    if (!B) return 0;
    int32_t n = 0;
    for (auto &b : B->inst) n += b.active ? 1 : 0;
    return n;
}

uint64_t gpe_batch_resident_bytes(const gpe_batch *B) { return B ? B->resident : 0; }

gpe_status gpe_batch_set_launch_mode(gpe_batch *B, uint32_t mode) {
    if (!B || mode > GPE_BATCH_STREAM) {
        set_error("gpe_batch_set_launch_mode: unknown mode");
        return GPE_EINVAL;
    }
    B->use_graph = mode == GPE_BATCH_GRAPH;
    B->graph_dirty = true;
    return GPE_OK;
}

gpe_status gpe_batch_set_group_bytes(gpe_batch *B, uint64_t bytes) {
    if (!B) {
        set_error("gpe_batch_set_group_bytes: null batch");
        return GPE_EINVAL;
    }
    B->group_bytes = bytes;
    B->desc_dirty = true;  // groups are formed with the descriptors
    return GPE_OK;
}

uint64_t gpe_batch_group_bytes(const gpe_batch *B) { return B ? B->group_bytes : 0; }

gpe_status gpe_batch_set_tile(gpe_batch *B, uint32_t tile) {
    const uint32_t waves = tile & 0xff;
    const bool quad = (tile & GPE_BATCH_TILE_QUAD) != 0;
    const bool waves_ok = quad ? (waves == 3 || waves == 6 || waves == 12)
                               : (waves == 1 || waves == 2 || waves == 4 || waves == 8);
    if (!B || !waves_ok ||
        (tile & ~(0xffu | GPE_BATCH_TILE_CACHED | GPE_BATCH_TILE_SLABS | GPE_BATCH_TILE_GROUP_SLABS |
                  GPE_BATCH_TILE_QUAD)) ||
        ((tile & GPE_BATCH_TILE_SLABS) && (tile & GPE_BATCH_TILE_GROUP_SLABS))) {
        set_error("gpe_batch_set_tile: waves must be 1, 2, 4 or 8, or 3, 6 or 12 with GPE_BATCH_TILE_QUAD "
                  "(optionally | GPE_BATCH_TILE_CACHED, | GPE_BATCH_TILE_SLABS or | GPE_BATCH_TILE_GROUP_SLABS)");
        return GPE_EINVAL;
    }
    for (auto &b : B->inst)
        if (b.active) {
            set_error("gpe_batch_set_tile: the batch holds instances");
            return GPE_EBUSY;
        }
    B->real_ty = (int)waves;
    B->real_quad = quad;
    B->real_nt = !(tile & GPE_BATCH_TILE_CACHED);
    B->real_order = (tile & GPE_BATCH_TILE_GROUP_SLABS) ? 2 : (tile & GPE_BATCH_TILE_SLABS) ? 1 : 0;
    B->graph_dirty = true;
    return GPE_OK;
}

}  // extern "C"
