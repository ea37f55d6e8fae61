// gpe_kernels.hpp -- gfx950 device code of the BCC Gross-Pitaevskii stepper.
//
// Field layout in HBM (one buffer per time level, "class-row SoA"):
//   element (i, j, k, l) at  origin + k*sz + j*sy + l*xp + i
//   with xp = row pitch (elements), sy = 12*xp, sz = (ny+2)*sy.
// For each cell row (j,k) the 12 class rows are contiguous, each holding the
// nx cells of that class in x order, so a 64-lane wavefront reading one
// class of 64 consecutive cells issues one fully coalesced 1 KiB (FP64) or
// 512 B (FP32) access.  A neighbour in another cell is the same class row
// shifted by dx + dy*sy + dz*sz: still coalesced, no gathers.  One zero halo
// cell surrounds the box on every side (homogeneous Dirichlet boundary, as
// the never-written halo of the reference's pitched buffers,
// GPUGrossPitaevskii_BCC.cu:269-287).  Rows start on 128-byte lines and
// xp >= nx + 1, so the x halo at x = nx is padding of its own row and the one
// at x = -1 is the last padding element of the row before it (one line in
// front of the first row holds that of the first): no line of a row holds
// only halo on the left, and the line that ends a row serves both halos.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bcc_stencil.hpp"

namespace gpe {

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };
template <typename T> using cplx_t = typename cplx<T>::type;

struct Layout {
    int nx, ny, nz;
    int line;            // elements per 128-byte line
    long long xp;        // class-row pitch in elements
    long long sy;        // 12*xp
    long long sz;        // (ny+2)*sy
    long long origin;    // element index of (0,0,0,l=0)
    long long elems;     // elements per field
};

// In-kernel trap V = 0.5*(x^2 + y^2 + rr*z^2) with node positions written
// in the operation order of GPUGrossPitaevskii_BCC.cu:258 (p0 + b*(i*sqrt(8)
// + q_l/sqrt(8))), so V*dt has the same bits as the host's potentialV3*dt.
struct Trap {
    double p0[3];
    double b;
    double bw;           // sqrt(8)
    double rr;           // ratio*ratio
    double bpos[12][3];  // q_l/sqrt(8)
    double q[4];         // (2n + 1)/sqrt(8): the four values bpos takes
};

// Node position of class l inside its cell, per axis, in units of b/sqrt(8)
// (1, 3, 5 or 7; GPUGrossPitaevskii_BCC.cu:76-88), as the index n of
// Trap::q (bpos[l][a] == q[class_q(l, a)]).
__host__ __device__ constexpr int class_q(int l, int a) {
    constexpr int pos[12][3] = {{1, 3, 7}, {7, 1, 3}, {7, 3, 1}, {7, 3, 5}, {7, 5, 3}, {1, 7, 3},
                                {3, 7, 1}, {3, 7, 5}, {5, 7, 3}, {3, 1, 7}, {3, 5, 7}, {5, 3, 7}};
    return pos[l][a] >> 1;
}

// Scalars of one solve.  *_dt are multiplied by the time step like the
// reference (GPUGrossPitaevskii_BCC.cu:304-307).
struct Coefs {
    double lapfac_dt, lapfac0_dt, g_dt, dt;
    double w;          // node measure sqrt(32/9) b^3
    double lap;        // 1.5/b^2 (unscaled Laplacian weight)
    double g;          // unscaled
};

enum { FORM_REF = 0, FORM_DIFF = 1 };   // neighbour-sum forms
enum { POT_ARRAY = 0, POT_TRAP = 1 };
enum { MODE_REAL = 0, MODE_IMAG = 1 };

__device__ __forceinline__ double trap_v(const Trap &t, int i, int j, int k, int l) {
#pragma clang fp contract(off)
    const double x = t.p0[0] + t.b * ((double)(unsigned)i * t.bw + t.bpos[l][0]);
    const double y = t.p0[1] + t.b * ((double)(unsigned)j * t.bw + t.bpos[l][1]);
    const double z = t.p0[2] + t.b * ((double)(unsigned)k * t.bw + t.bpos[l][2]);
    return 0.5 * (x * x + y * y + t.rr * z * z);
}

// The same value from the 4 distinct offsets.  Kernels with many live
// values use this form: the 36 offsets of bpos take 72 scalar registers.
// The one-thread-per-cell kernels keep bpos, where the compiler's sharing of
// the 4 offsets across classes costs vector registers and occupancy.
__device__ __forceinline__ double trap_v4(const Trap &t, int i, int j, int k, int l) {
#pragma clang fp contract(off)
    const double x = t.p0[0] + t.b * ((double)(unsigned)i * t.bw + t.q[class_q(l, 0)]);
    const double y = t.p0[1] + t.b * ((double)(unsigned)j * t.bw + t.q[class_q(l, 1)]);
    const double z = t.p0[2] + t.b * ((double)(unsigned)k * t.bw + t.q[class_q(l, 2)]);
    return 0.5 * (x * x + y * y + t.rr * z * z);
}

// V of a node in a one-thread-per-cell kernel computing in T (the same bits
// either way): the FP64 real-time update takes the per-class offsets, which
// leave it at 5 waves per SIMD; FP32 the 4-offset form, 1.4 % faster there at
// 128^3 cells; the FP64 imaginary-time update the 4-offset form too, which
// holds it at 3 waves per SIMD (133 registers): at 4 waves its blocks keep
// more layers in flight than an XCD's L2 holds, and it read 1.6 x as many
// bytes (704 against 442 MB per step at 128^3 cells) and ran 10 % slower.
template <typename T, int MODE = MODE_REAL>
__device__ __forceinline__ double trap_vt(const Trap &t, int i, int j, int k, int l) {
    if constexpr (sizeof(T) == 4 || MODE == MODE_IMAG) return trap_v4(t, i, j, k, l);
    else return trap_v(t, i, j, k, l);
}

// s = dt*H(psi)psi from the four neighbour values, |psi|^2 and V*dt.
//  FORM_REF  (FP64 parity): lapfac*((n0+n1)+n2)+n3) + ((lapfac0+pot)+g|psi|^2)psi,
//            the exact sequence of GPUGrossPitaevskii_BCC.cu:207-213.
//  FORM_DIFF (FP32 bulk):   lapfac*(sum_f (n_f - psi)) + (pot + g|psi|^2)psi,
//            algebraically equal (lapfac0 = -4 lapfac) and free of the
//            cancellation between lapfac*sum and lapfac0*psi.
template <typename T, int FORM>
struct Hamiltonian {
    using C = cplx_t<T>;
    T lapfac, lapfac0, g;
    __device__ __forceinline__ C apply(const C n[4], C psi, T pot) const {
#pragma clang fp contract(off)
        const T nsq = psi.x * psi.x + psi.y * psi.y;
        C s;
        if (FORM == FORM_REF) {
            T sx = n[0].x + n[1].x;
            T sy = n[0].y + n[1].y;
            sx = sx + n[2].x; sy = sy + n[2].y;
            sx = sx + n[3].x; sy = sy + n[3].y;
            const T c = (lapfac0 + pot) + g * nsq;
            s.x = lapfac * sx + c * psi.x;
            s.y = lapfac * sy + c * psi.y;
        } else {
            T dx = (n[0].x - psi.x) + (n[1].x - psi.x);
            T dy = (n[0].y - psi.y) + (n[1].y - psi.y);
            dx = dx + (n[2].x - psi.x); dy = dy + (n[2].y - psi.y);
            dx = dx + (n[3].x - psi.x); dy = dy + (n[3].y - psi.y);
            const T c = pot + g * nsq;
            s.x = lapfac * dx + c * psi.x;
            s.y = lapfac * dy + c * psi.y;
        }
        return s;
    }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Deterministic block sum of NV FP64 values; result valid in thread 0.
template <int NV, int NWAVES>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds /* NV*NWAVES */) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) lds[q * NWAVES + wave] = v[q];
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double s = 0.0;
            for (int w = 0; w < NWAVES; w++) s += lds[q * NWAVES + w];
            v[q] = s;
        }
    }
}

template <typename T>
__device__ __forceinline__ cplx_t<T> cscale(cplx_t<T> a, T s) {
    cplx_t<T> r;
    r.x = s * a.x;
    r.y = s * a.y;
    return r;
}

// Logical block coordinates from a 1-D launch.  Workgroups are dispatched
// round-robin over the 8 XCDs (each with its own L2): block lin runs on XCD
// lin%8.  remap 0: block lin is tile lin of the (bx fastest, then by, then
// bz) order.  remap 1: XCD x takes the x-th contiguous eighth of that order,
// so blocks sharing halo rows and layers share one L2 (a bijection for any
// block count).  remap 2 (gy a multiple of 8): XCD x owns the y-band
// [x gy/8, (x+1) gy/8) of every layer and walks it layer by layer, so y-halo
// rows and z-neighbours stay in one L2 while all XCDs sweep z together.
// rev walks the layers (remap 2) or the whole order (remap 0, 1) backwards.
struct Grid3 {
    int gx, gy, gz, remap;
    int lsh;  // log2 of the lanes one x-row of cells takes (6: a wave per row; 5, 4: 2 or 4 rows per wave)
    int rev;  // 1: tiles in descending order (the pass starts where the previous one ended)
    int cz;   // layers per z-chunk of z-march kernels that take it at launch (0: compile-time chunks)
};

// Lanes per x-row for an nx-cell grid: the width in {64, 32, 16} with the
// least padding (ties to the wider row), so narrow grids keep lanes busy.
__host__ __device__ inline int row_lanes_log2(int nx) {
    int best = 6, waste = ((nx + 63) / 64) * 64 - nx;
    for (int s = 5; s >= 4; s--) {
        const int w = 1 << s, pad = ((nx + w - 1) / w) * w - nx;
        if (pad < waste) {
            waste = pad;
            best = s;
        }
    }
    return best;
}
__device__ __forceinline__ void block_coords(const Grid3 &G, int &bx, int &by, int &bz) {
    const int nb = G.gx * G.gy * G.gz;
    const int lin = blockIdx.x;
    if (G.remap == 2) {
        const int br = G.gy >> 3, bs = G.gx * br, xcd = lin & 7, slot = lin >> 3;
        const int layer = slot / bs, within = slot - layer * bs;
        bz = G.rev ? G.gz - 1 - layer : layer;
        by = xcd * br + within / G.gx;
        bx = within % G.gx;
        return;
    }
    int w = lin;
    if (G.remap) {
        const int q = nb >> 3, r = nb & 7, xcd = lin & 7, slot = lin >> 3;
        w = (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
    }
    if (G.rev) w = nb - 1 - w;
    bx = w % G.gx;
    const int t = w / G.gx;
    by = t % G.gy;
    bz = t / G.gy;
}

// ---------------------------------------------------------------------
// GATHER: one thread per cell, all 12 classes; in-cell neighbours come from
// the thread's own registers, the 18 off-cell neighbours per cell from
// global memory (cache hits for all but first touch).  Logical grid
// (ceil(nx/64), ceil(ny/TY), nz), block (64, TY).
// Real time: next[l] += (s.y, -s.x) in place (GPUGrossPitaevskii_BCC.cu:215).
// Imaginary time: out = psi - s, psi = scale*prev, block partial of
// sum|out|^2 to partials[block].
// ---------------------------------------------------------------------
// Cell (bx*64 + threadIdx.x, by*TY + threadIdx.y, bz) of the gather pass for
// the calling thread; returns its sum of |out|^2 (imaginary time).  NT
// marks the read-modify-write operand as streaming (read once, written
// once) so it does not displace neighbour rows from the caches.
// Unscaled weights of the observables (k_observables): 1/dt turns a
// streamed V*dt back into V, lap = 1.5/b^2, g.
struct ObsCoefs {
    double inv_dt, lap, g;
};

// OBS (real time): also accumulate into obs[4] the observables N, K, P, I
// (k_observables' per-node terms, FP64) of the field being read, from the
// values the update already holds -- no extra memory traffic.
template <typename T, int FORM, int POT, int MODE, int TY, bool NT = false, bool OBS = false>
__device__ __forceinline__ double gather_tile(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                              const T *__restrict__ pot, const Layout &L,
                                              const Hamiltonian<T, FORM> &H, const Trap &trap, double dt,
                                              const double *__restrict__ scale, int bx, int by, int bz,
                                              int lsh = 6, double *obs = nullptr, ObsCoefs oc = {}) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    // a wave covers 64 >> lsh rows of (1 << lsh) cells
    const int i = (bx << lsh) + (threadIdx.x & ((1 << lsh) - 1));
    const int j = ((by * TY + threadIdx.y) << (6 - lsh)) + (threadIdx.x >> lsh);
    const int k = bz;
    const bool act = i < L.nx && j < L.ny;
    double part[1] = {0.0};
    if (act) {
        const long long base = L.origin + (long long)k * L.sz + (long long)j * L.sy + i;
        T s = (T)1;
        if (MODE == MODE_IMAG) s = (T)(*scale);
        C own[12];
#pragma unroll
        for (int l = 0; l < 12; l++) {
            own[l] = prev[base + l * L.xp];
            if (MODE == MODE_IMAG) own[l] = cscale<T>(own[l], s);
        }
#pragma unroll
        for (int l = 0; l < 12; l++) {
            C n[4];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const Face e = bcc_face(l, f);
                if (face_in_cell(e)) {
                    n[f] = own[e.m];
                } else {
                    n[f] = prev[base + e.m * L.xp + e.dx + e.dy * L.sy + e.dz * L.sz];
                    if (MODE == MODE_IMAG) n[f] = cscale<T>(n[f], s);
                }
            }
            T pv;
            if (POT == POT_TRAP) pv = (T)(trap_vt<T, MODE>(trap, i, j, k, l) * dt);
            else pv = pot[base + l * L.xp];
            if (OBS) {
                const double pr = (double)own[l].x, pi = (double)own[l].y;
                double dxs = 0.0, dys = 0.0;
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    dxs += (double)n[f].x - pr;
                    dys += (double)n[f].y - pi;
                }
                const double n2 = pr * pr + pi * pi;
                const double v = POT == POT_TRAP ? trap_v(trap, i, j, k, l) : (double)pv * oc.inv_dt;
                obs[0] += n2;
                obs[1] += -0.5 * oc.lap * (pr * dxs + pi * dys);
                obs[2] += v * n2;
                obs[3] += oc.g * n2 * n2;
            }
            const C sres = H.apply(n, own[l], pv);
            const long long at = base + l * L.xp;
            if (MODE == MODE_REAL) {
                C nx;
                if (NT) {
                    nx.x = __builtin_nontemporal_load(&next[at].x);
                    nx.y = __builtin_nontemporal_load(&next[at].y);
                } else {
                    nx = next[at];
                }
                nx.x = nx.x + sres.y;
                nx.y = nx.y + (-sres.x);
                if (NT) {
                    __builtin_nontemporal_store(nx.x, &next[at].x);
                    __builtin_nontemporal_store(nx.y, &next[at].y);
                } else {
                    next[at] = nx;
                }
            } else {
                C o;
                o.x = own[l].x - sres.x;
                o.y = own[l].y - sres.y;
                // written once, read only by the next step: streaming store,
                // or a cached one where the fields fit the cache (NT false)
                if (NT) {
                    __builtin_nontemporal_store(o.x, &next[at].x);
                    __builtin_nontemporal_store(o.y, &next[at].y);
                } else {
                    next[at] = o;
                }
                part[0] += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
            }
        }
    }
    return part[0];
}

template <typename T, int FORM, int POT, int MODE, int TY, bool NT = false>
__global__ void __launch_bounds__(64 * TY) k_gather(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                                    const T *__restrict__ pot, Layout L, Hamiltonian<T, FORM> H,
                                                    Trap trap, double dt, const double *__restrict__ scale,
                                                    double *__restrict__ partials, Grid3 G) {
    __shared__ double lds[TY];
    int bx, by, bz;
    block_coords(G, bx, by, bz);
    double part[1];
    part[0] = gather_tile<T, FORM, POT, MODE, TY, NT>(next, prev, pot, L, H, trap, dt, scale, bx, by, bz, G.lsh);
    if (MODE == MODE_IMAG) {
        // partials in logical tile order (independent of the block-to-XCD order)
        block_sum<1, TY>(part, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[bx + G.gx * (by + G.gy * bz)] = part[0];
    }
}

// Real-time gather pass that also writes block partials of the observables
// of `prev` (4 per block, logical tile order) for k_finalize<4>.
template <typename T, int FORM, int POT, int TY, bool NT>
__global__ void __launch_bounds__(64 * TY)
    k_gather_obs(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev, const T *__restrict__ pot,
                 Layout L, Hamiltonian<T, FORM> H, Trap trap, double dt, ObsCoefs oc,
                 double *__restrict__ partials, Grid3 G) {
    __shared__ double lds[4 * TY];
    int bx, by, bz;
    block_coords(G, bx, by, bz);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    gather_tile<T, FORM, POT, MODE_REAL, TY, NT, true>(next, prev, pot, L, H, trap, dt, nullptr, bx, by, bz, G.lsh,
                                                       acc, oc);
    block_sum<4, TY>(acc, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const long long b = bx + (long long)G.gx * (by + (long long)G.gy * bz);
#pragma unroll
        for (int q = 0; q < 4; q++) partials[4 * b + q] = acc[q];
    }
}

// ---------------------------------------------------------------------
// NODE: one thread per node (i, j, k, l), the reference's own granularity
// (GPUGrossPitaevskii_BCC.cu:176-216), for grids too small to keep the chip
// busy with one thread per cell (32^3 cells make 512 one-wave gather blocks).
// A wave holds one class l of 64 >> lsh cell rows; l is wave-uniform, so
// each of the 12 code paths has its faces and trap offsets at compile time.
// A block of TY waves holds TY consecutive classes of the same cell rows
// (TY = 12: whole cells, whose 12 class rows then share the CU's L1).  Every
// node reads its own value and its 4 neighbours; the arithmetic per node is
// the gather pass's, so the bits are the same.  Logical waves: 12 x gx x gy x
// nz with gx = ceil(nx / (1 << lsh)), gy = ceil(ny / (64 >> lsh)), classes
// fastest; blocks ceil(that / TY).
// ---------------------------------------------------------------------
template <typename T, int FORM, int POT, int MODE, int CL, bool NT>
__device__ __forceinline__ double node_update(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                              const T *__restrict__ pot, const Layout &L,
                                              const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, T s,
                                              int i, int j, int k) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    const long long base = L.origin + (long long)k * L.sz + (long long)j * L.sy + i;
    const long long at = base + CL * L.xp;
    C own = prev[at];
    if (MODE == MODE_IMAG) own = cscale<T>(own, s);
    C n[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const Face e = bcc_face(CL, f);
        if (face_in_cell(e)) n[f] = prev[base + e.m * L.xp];
        else n[f] = prev[base + e.m * L.xp + e.dx + e.dy * L.sy + e.dz * L.sz];
        if (MODE == MODE_IMAG) n[f] = cscale<T>(n[f], s);
    }
    T pv;
    if (POT == POT_TRAP) pv = (T)(trap_vt<T, MODE>(trap, i, j, k, CL) * dt);
    else pv = pot[at];
    const C sres = H.apply(n, own, pv);
    if (MODE == MODE_REAL) {
        C nx;
        if (NT) {
            nx.x = __builtin_nontemporal_load(&next[at].x);
            nx.y = __builtin_nontemporal_load(&next[at].y);
        } else {
            nx = next[at];
        }
        nx.x = nx.x + sres.y;
        nx.y = nx.y + (-sres.x);
        if (NT) {
            __builtin_nontemporal_store(nx.x, &next[at].x);
            __builtin_nontemporal_store(nx.y, &next[at].y);
        } else {
            next[at] = nx;
        }
        return 0.0;
    } else {
        C o;
        o.x = own.x - sres.x;
        o.y = own.y - sres.y;
        if (NT) {
            __builtin_nontemporal_store(o.x, &next[at].x);
            __builtin_nontemporal_store(o.y, &next[at].y);
        } else {
            next[at] = o;
        }
        return (double)o.x * (double)o.x + (double)o.y * (double)o.y;
    }
}

// Imaginary time: block partials of sum|out|^2 in logical block order.
// G.remap 1: XCD x takes the x-th contiguous eighth of the blocks (as
// block_coords); G.rev walks the blocks backwards.
template <typename T, int FORM, int POT, int MODE, int TY, bool NT = false>
__global__ void __launch_bounds__(64 * TY) k_node(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                                  const T *__restrict__ pot, Layout L, Hamiltonian<T, FORM> H,
                                                  Trap trap, double dt, const double *__restrict__ scale,
                                                  double *__restrict__ partials, Grid3 G) {
    __shared__ double lds[TY];
    const int nb = (int)gridDim.x, lin = (int)blockIdx.x;
    int b = lin;
    if (G.remap == 1) {
        const int q = nb >> 3, r = nb & 7, xcd = lin & 7, slot = lin >> 3;
        b = (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
    }
    if (G.rev) b = nb - 1 - b;
    const int wv = b * TY + (int)threadIdx.y;  // logical wave
    const int cl = __builtin_amdgcn_readfirstlane(wv % 12);
    int t = wv / 12;
    const int bx = t % G.gx;
    t /= G.gx;
    const int by = t % G.gy, k = t / G.gy;
    const int i = (bx << G.lsh) + ((int)threadIdx.x & ((1 << G.lsh) - 1));
    const int j = (by << (6 - G.lsh)) + ((int)threadIdx.x >> G.lsh);
    double part = 0.0;
    if (k < L.nz && i < L.nx && j < L.ny) {
        T s = (T)1;
        if (MODE == MODE_IMAG) s = (T)(*scale);
#define GPE_NODE_CASE(C_)                                                                                 \
    case C_:                                                                                              \
        part = node_update<T, FORM, POT, MODE, C_, NT>(next, prev, pot, L, H, trap, dt, s, i, j, k);      \
        break;
        switch (cl) {
            GPE_NODE_CASE(0)
            GPE_NODE_CASE(1)
            GPE_NODE_CASE(2)
            GPE_NODE_CASE(3)
            GPE_NODE_CASE(4)
            GPE_NODE_CASE(5)
            GPE_NODE_CASE(6)
            GPE_NODE_CASE(7)
            GPE_NODE_CASE(8)
            GPE_NODE_CASE(9)
            GPE_NODE_CASE(10)
            default: GPE_NODE_CASE(11)
        }
#undef GPE_NODE_CASE
    }
    if (MODE == MODE_IMAG) {
        double p[1] = {part};
        block_sum<1, TY>(p, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[b] = p[0];
    }
}

// ---------------------------------------------------------------------
// QUAD: one thread per cell and group of 4 classes, between the gather pass
// (12 classes per thread) and the node pass (1): three times the waves of the
// gather pass for about the same loads.  The groups {0,9,10,11}, {1,2,3,4},
// {5,6,7,8} keep most in-cell neighbours inside the thread's own registers;
// the others are loads (L1 hits).  A wave holds one group (wave-uniform) of
// 64 >> lsh cell rows; a block of TY waves holds TY consecutive (group, rows)
// pairs, groups fastest.  Per node the arithmetic is the gather pass's.
// Logical waves 3 x gx x gy x nz, blocks ceil(that / TY).
// ---------------------------------------------------------------------
__host__ __device__ constexpr int quad_class(int grp, int q) {
    return grp == 0 ? (q == 0 ? 0 : 8 + q) : 4 * grp - 3 + q;
}
__host__ __device__ constexpr int quad_slot(int grp, int m) {
    for (int q = 0; q < 4; q++)
        if (quad_class(grp, q) == m) return q;
    return -1;
}

template <typename T, int FORM, int POT, int MODE, int GRP, bool NT>
__device__ __forceinline__ double quad_update(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                              const T *__restrict__ pot, const Layout &L,
                                              const Hamiltonian<T, FORM> &H, const Trap &trap, double dt, T s,
                                              int i, int j, int k) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    const long long base = L.origin + (long long)k * L.sz + (long long)j * L.sy + i;
    C own[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        own[q] = prev[base + quad_class(GRP, q) * L.xp];
        if (MODE == MODE_IMAG) own[q] = cscale<T>(own[q], s);
    }
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int l = quad_class(GRP, q);
        C n[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            const Face e = bcc_face(l, f);
            if (face_in_cell(e) && quad_slot(GRP, e.m) >= 0) {
                n[f] = own[quad_slot(GRP, e.m)];
            } else {
                n[f] = face_in_cell(e) ? prev[base + e.m * L.xp]
                                       : prev[base + e.m * L.xp + e.dx + e.dy * L.sy + e.dz * L.sz];
                if (MODE == MODE_IMAG) n[f] = cscale<T>(n[f], s);
            }
        }
        const long long at = base + l * L.xp;
        T pv;
        if (POT == POT_TRAP) pv = (T)(trap_vt<T, MODE>(trap, i, j, k, l) * dt);
        else pv = pot[at];
        const C sres = H.apply(n, own[q], pv);
        if (MODE == MODE_REAL) {
            C nx;
            if (NT) {
                nx.x = __builtin_nontemporal_load(&next[at].x);
                nx.y = __builtin_nontemporal_load(&next[at].y);
            } else {
                nx = next[at];
            }
            nx.x = nx.x + sres.y;
            nx.y = nx.y + (-sres.x);
            if (NT) {
                __builtin_nontemporal_store(nx.x, &next[at].x);
                __builtin_nontemporal_store(nx.y, &next[at].y);
            } else {
                next[at] = nx;
            }
        } else {
            C o;
            o.x = own[q].x - sres.x;
            o.y = own[q].y - sres.y;
            if (NT) {
                __builtin_nontemporal_store(o.x, &next[at].x);
                __builtin_nontemporal_store(o.y, &next[at].y);
            } else {
                next[at] = o;
            }
            part += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
        }
    }
    return part;
}

// Imaginary time: block partials of sum|out|^2 in logical block order.
// G.remap 1: XCD x takes the x-th contiguous eighth of the blocks; 2: y-bands
// per XCD; G.rev walks the blocks (the layers) backwards.
template <typename T, int FORM, int POT, int MODE, int TY, bool NT = false>
__global__ void __launch_bounds__(64 * TY) k_quad(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                                  const T *__restrict__ pot, Layout L, Hamiltonian<T, FORM> H,
                                                  Trap trap, double dt, const double *__restrict__ scale,
                                                  double *__restrict__ partials, Grid3 G) {
    __shared__ double lds[TY];
    const int nb = (int)gridDim.x, lin = (int)blockIdx.x;
    int b = lin;
    if (G.remap == 2) {
        // XCD x owns the y-band [x gy/8, (x+1) gy/8) of wave rows of every
        // layer and walks it layer by layer (G.rev: layers backwards); the
        // band of one layer is bb whole blocks (the host checks gy % 8 == 0
        // and 3 gx gy/8 % TY == 0)
        const int bb = 3 * G.gx * (G.gy >> 3) / TY, xcd = lin & 7, slot = lin >> 3;
        const int layer = slot / bb, within = slot - layer * bb;
        b = ((G.rev ? G.gz - 1 - layer : layer) * 8 + xcd) * bb + within;
    } else {
        if (G.remap == 1) {
            const int q = nb >> 3, r = nb & 7, xcd = lin & 7, slot = lin >> 3;
            b = (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
        }
        if (G.rev) b = nb - 1 - b;
    }
    const int wv = b * TY + (int)threadIdx.y;  // logical wave
    const int grp = __builtin_amdgcn_readfirstlane(wv % 3);
    int t = wv / 3;
    const int bx = t % G.gx;
    t /= G.gx;
    const int by = t % G.gy, k = t / G.gy;
    const int i = (bx << G.lsh) + ((int)threadIdx.x & ((1 << G.lsh) - 1));
    const int j = (by << (6 - G.lsh)) + ((int)threadIdx.x >> G.lsh);
    double part = 0.0;
    if (k < L.nz && i < L.nx && j < L.ny) {
        T s = (T)1;
        if (MODE == MODE_IMAG) s = (T)(*scale);
        if (grp == 0) part = quad_update<T, FORM, POT, MODE, 0, NT>(next, prev, pot, L, H, trap, dt, s, i, j, k);
        else if (grp == 1) part = quad_update<T, FORM, POT, MODE, 1, NT>(next, prev, pot, L, H, trap, dt, s, i, j, k);
        else part = quad_update<T, FORM, POT, MODE, 2, NT>(next, prev, pot, L, H, trap, dt, s, i, j, k);
    }
    if (MODE == MODE_IMAG) {
        double p[1] = {part};
        block_sum<1, TY>(p, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[b] = p[0];
    }
}

// ---------------------------------------------------------------------
// ZMARCH: one thread per (i,j) column walking CZ layers in z.  The thread
// keeps its own cell's 12 values, classes {10,11} of the layer below and
// classes {2,6} of the layer above in registers -- exactly the values the
// (0,0,-1) and (0,0,+1) faces need -- so z-neighbours never touch memory
// twice; x/y/diagonal neighbours are global loads that hit L1/L2.
// Logical grid (ceil(nx/64), ceil(ny/TY), ceil(nz/CZ)), block (64, TY).
// ---------------------------------------------------------------------
// One (bx, by, bz) tile of the z-march for the calling thread; returns the
// thread's sum of |out|^2 (imaginary time) for the block reduction.
template <typename T, int FORM, int POT, int MODE, int TY, bool NT = false>
__device__ __forceinline__ double zmarch_tile(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                              const T *__restrict__ pot, const Layout &L,
                                              const Hamiltonian<T, FORM> &H, const Trap &trap, double dt,
                                              const double *__restrict__ scale, int bx, int by, int bz, int cz,
                                              int lsh = 6) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    const int i = (bx << lsh) + (threadIdx.x & ((1 << lsh) - 1));
    const int j = ((by * TY + threadIdx.y) << (6 - lsh)) + (threadIdx.x >> lsh);
    const int k0 = bz * cz;
    const int k1 = min(k0 + cz, L.nz);
    const bool act = i < L.nx && j < L.ny;
    double part[1] = {0.0};
    if (act) {
        T s = (T)1;
        if (MODE == MODE_IMAG) s = (T)(*scale);
        const long long col = L.origin + (long long)j * L.sy + i;
        C cur[12], dn10, dn11;
        {
            const long long b0 = col + (long long)k0 * L.sz;
#pragma unroll
            for (int l = 0; l < 12; l++) cur[l] = prev[b0 + l * L.xp];
            dn10 = prev[b0 - L.sz + 10 * L.xp];
            dn11 = prev[b0 - L.sz + 11 * L.xp];
            if (MODE == MODE_IMAG) {
#pragma unroll
                for (int l = 0; l < 12; l++) cur[l] = cscale<T>(cur[l], s);
                dn10 = cscale<T>(dn10, s);
                dn11 = cscale<T>(dn11, s);
            }
        }
        for (int k = k0; k < k1; k++) {
            const long long base = col + (long long)k * L.sz;
            C up2 = prev[base + L.sz + 2 * L.xp];
            C up6 = prev[base + L.sz + 6 * L.xp];
            if (MODE == MODE_IMAG) {
                up2 = cscale<T>(up2, s);
                up6 = cscale<T>(up6, s);
            }
#pragma unroll
            for (int l = 0; l < 12; l++) {
                C n[4];
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    const Face e = bcc_face(l, f);
                    if (face_in_cell(e)) {
                        n[f] = cur[e.m];
                    } else if (e.dx == 0 && e.dy == 0 && e.dz == -1) {
                        n[f] = (e.m == 10) ? dn10 : dn11;
                    } else if (e.dx == 0 && e.dy == 0 && e.dz == 1) {
                        n[f] = (e.m == 2) ? up2 : up6;
                    } else {
                        n[f] = prev[base + e.m * L.xp + e.dx + e.dy * L.sy + e.dz * L.sz];
                        if (MODE == MODE_IMAG) n[f] = cscale<T>(n[f], s);
                    }
                }
                T pv;
                if (POT == POT_TRAP) pv = (T)(trap_vt<T, MODE>(trap, i, j, k, l) * dt);
                else pv = pot[base + l * L.xp];
                const C sres = H.apply(n, cur[l], pv);
                const long long at = base + l * L.xp;
                if (MODE == MODE_REAL) {
                    C nx;
                    if (NT) {
                        nx.x = __builtin_nontemporal_load(&next[at].x);
                        nx.y = __builtin_nontemporal_load(&next[at].y);
                    } else {
                        nx = next[at];
                    }
                    nx.x = nx.x + sres.y;
                    nx.y = nx.y + (-sres.x);
                    if (NT) {
                        __builtin_nontemporal_store(nx.x, &next[at].x);
                        __builtin_nontemporal_store(nx.y, &next[at].y);
                    } else {
                        next[at] = nx;
                    }
                } else {
                    C o;
                    o.x = cur[l].x - sres.x;
                    o.y = cur[l].y - sres.y;
                    // written once, read only by the next step: streaming store
                    // (NT), or a cached one
                    if (NT) {
                        __builtin_nontemporal_store(o.x, &next[at].x);
                        __builtin_nontemporal_store(o.y, &next[at].y);
                    } else {
                        next[at] = o;
                    }
                    part[0] += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
                }
            }
            dn10 = cur[10];
            dn11 = cur[11];
            if (k + 1 < k1) {
                const long long nb = base + L.sz;
#pragma unroll
                for (int l = 0; l < 12; l++) {
                    if (l == 2) cur[l] = up2;
                    else if (l == 6) cur[l] = up6;
                    else {
                        cur[l] = prev[nb + l * L.xp];
                        if (MODE == MODE_IMAG) cur[l] = cscale<T>(cur[l], s);
                    }
                }
            }
        }
    }
    return part[0];
}

template <typename T, int FORM, int POT, int MODE, int TY, int CZ, bool NT = false>
__global__ void __launch_bounds__(64 * TY) k_zmarch(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                                    const T *__restrict__ pot, Layout L, Hamiltonian<T, FORM> H,
                                                    Trap trap, double dt, const double *__restrict__ scale,
                                                    double *__restrict__ partials, Grid3 G) {
    __shared__ double lds[TY];
    int bx, by, bz;
    block_coords(G, bx, by, bz);
    double part[1];
    part[0] = zmarch_tile<T, FORM, POT, MODE, TY, NT>(next, prev, pot, L, H, trap, dt, scale, bx, by, bz, CZ, G.lsh);
    if (MODE == MODE_IMAG) {
        // partials in logical tile order, so the norm's summation order does
        // not depend on the block-to-XCD order
        block_sum<1, TY>(part, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[bx + G.gx * (by + G.gy * bz)] = part[0];
    }
}

// ---------------------------------------------------------------------
// Observables of one field: per-block FP64 partial sums of N, K, P, I
// (see vaikkari_gpe.h), written to partials[4*block + q].  FP32 fields are
// widened to FP64 before any arithmetic.
// ---------------------------------------------------------------------
template <typename T, int POT, int TY>
__global__ void __launch_bounds__(64 * TY) k_observables(const cplx_t<T> *__restrict__ psi, const T *__restrict__ pot,
                                                         Layout L, Trap trap, double inv_dt, double lap, double g,
                                                         double *__restrict__ partials) {
    using C = cplx_t<T>;
    __shared__ double lds[4 * TY];
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int k = blockIdx.z;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < L.nx && j < L.ny) {
        const long long base = L.origin + (long long)k * L.sz + (long long)j * L.sy + i;
#pragma unroll
        for (int l = 0; l < 12; l++) {
            const C c = psi[base + l * L.xp];
            const double pr = (double)c.x, pi = (double)c.y;
            double dx = 0.0, dy = 0.0;
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const Face e = bcc_face(l, f);
                const C nb = psi[base + e.m * L.xp + e.dx + e.dy * L.sy + e.dz * L.sz];
                dx += (double)nb.x - pr;
                dy += (double)nb.y - pi;
            }
            const double n2 = pr * pr + pi * pi;
            double v;
            if (POT == POT_TRAP) v = trap_v(trap, i, j, k, l);
            else v = (double)pot[base + l * L.xp] * inv_dt;
            acc[0] += n2;
            acc[1] += -0.5 * lap * (pr * dx + pi * dy);
            acc[2] += v * n2;
            acc[3] += g * n2 * n2;
        }
    }
    block_sum<4, TY>(acc, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
#pragma unroll
        for (int q = 0; q < 4; q++) partials[4 * b + q] = acc[q];
    }
}

// Distance of two fields on the same cell grid, each in its own layout and
// precision: per block {sum |a - b|^2, sum |b|^2, sum |a|^2} over its nodes,
// in FP64, for k_finalize<3>.
template <typename TA, typename TB, int TY>
__global__ void __launch_bounds__(64 * TY) k_field_distance(const cplx_t<TA> *__restrict__ a, Layout La,
                                                            const cplx_t<TB> *__restrict__ b, Layout Lb,
                                                            double *__restrict__ partials) {
    __shared__ double lds[3 * TY];
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int k = blockIdx.z;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i < La.nx && j < La.ny) {
        const long long ba = La.origin + (long long)k * La.sz + (long long)j * La.sy + i;
        const long long bb = Lb.origin + (long long)k * Lb.sz + (long long)j * Lb.sy + i;
#pragma unroll
        for (int l = 0; l < 12; l++) {
            const cplx_t<TA> u = a[ba + l * La.xp];
            const cplx_t<TB> v = b[bb + l * Lb.xp];
            const double ur = (double)u.x, ui = (double)u.y, vr = (double)v.x, vi = (double)v.y;
            const double dr = ur - vr, di = ui - vi;
            acc[0] += dr * dr + di * di;
            acc[1] += vr * vr + vi * vi;
            acc[2] += ur * ur + ui * ui;
        }
    }
    block_sum<3, TY>(acc, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const long long bl = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
#pragma unroll
        for (int q = 0; q < 3; q++) partials[3 * bl + q] = acc[q];
    }
}

// One block of 256 threads reduces NV interleaved partial streams in a
// fixed order (Neumaier inside each thread, then a fixed tree), multiplies by
// w and writes out[q].  For the imaginary-time norm (NV == 1, renorm != 0)
// it also writes out[1] = 1/sqrt(N), the scale the next step applies to
// every value it reads.
// Neumaier-compensated sum of p[b * pitch] for b = first, first + step, ...
// below n, added in that order.  The loads go out eight at a time ahead of
// their additions (one memory latency per eight terms instead of per term);
// the order of the additions, and so the result, is that of a plain loop.
__device__ __forceinline__ void neumaier_add(double &s, double &c, double x) {
#pragma clang fp contract(off)
    const double u = s + x;
    c += (fabs(s) >= fabs(x)) ? ((s - u) + x) : ((x - u) + s);
    s = u;
}
__device__ __forceinline__ double strided_neumaier(const double *__restrict__ p, long long pitch, long long first,
                                                   long long step, long long n) {
#pragma clang fp contract(off)
    double s = 0.0, c = 0.0;
    long long b = first;
    for (; b + 7 * step < n; b += 8 * step) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = p[(b + u * step) * pitch];
#pragma unroll
        for (int u = 0; u < 8; u++) neumaier_add(s, c, x[u]);
    }
    for (; b < n; b += step) neumaier_add(s, c, p[b * pitch]);
    return s + c;
}

template <int NV>
__global__ void __launch_bounds__(256) k_finalize(const double *__restrict__ partials, long long nblocks, double w,
                                                  int renorm, double *__restrict__ out) {
    __shared__ double red[NV][256];
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < NV; q++) red[q][t] = strided_neumaier(partials + q, NV, t, 256, nblocks);
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int q = 0; q < NV; q++) red[q][t] += red[q][t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) out[q] = w * red[q][0];
        if (renorm) out[1] = 1.0 / sqrt(out[0]);
    }
}

// a *= s elementwise over a whole buffer (halo zeros stay zero).
template <typename T>
__global__ void k_scale_field(cplx_t<T> *__restrict__ a, long long n, const double *__restrict__ scale) {
    const T s = (T)(*scale);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        cplx_t<T> v = a[e];
        v.x = s * v.x;
        v.y = s * v.y;
        a[e] = v;
    }
}

// Counter-based N(0,1) pair for node n: splitmix64 of (seed, n), Box-Muller.
// The Gaussian recipe of batched, served and device-initialised solves:
// even = amp exp(-r^2 inv2s2) (1 + noise (xr + i xi)) at every node, odd (if
// given) = even (cph + i sph).
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

template <typename T>
__global__ void k_init_gaussian(cplx_t<T> *__restrict__ even, cplx_t<T> *__restrict__ odd, Layout L, Trap trap,
                                double inv2s2, double amp, double noise, unsigned long long seed, double cph,
                                double sph) {
    const long long cells = (long long)L.nx * L.ny * L.nz;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < cells;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % L.nx), j = (int)((c / L.nx) % L.ny), k = (int)(c / ((long long)L.nx * L.ny));
        const long long base = L.origin + (long long)k * L.sz + (long long)j * L.sy + i;
        for (int l = 0; l < 12; l++) {
            const double x = trap.p0[0] + trap.b * ((double)(unsigned)i * trap.bw + trap.bpos[l][0]);
            const double y = trap.p0[1] + trap.b * ((double)(unsigned)j * trap.bw + trap.bpos[l][1]);
            const double z = trap.p0[2] + trap.b * ((double)(unsigned)k * trap.bw + trap.bpos[l][2]);
            const double a = amp * exp(-(x * x + y * y + z * z) * inv2s2);
            const unsigned long long h1 = splitmix64(seed ^ (unsigned long long)(c * 12 + l) * 0xD1B54A32D192ED03ULL);
            const unsigned long long h2 = splitmix64(h1);
            const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
            const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
            const double r = sqrt(-2.0 * log(u1));
            const double xr = r * cos(6.2831853071795864770 * u2), xi = r * sin(6.2831853071795864770 * u2);
            const double er = a * (1.0 + noise * xr), ei = a * (noise * xi);
            cplx_t<T> e, o;
            e.x = (T)er;
            e.y = (T)ei;
            even[base + l * L.xp] = e;
            if (odd) {
                o.x = (T)(cph * er - sph * ei);
                o.y = (T)(sph * er + cph * ei);
                odd[base + l * L.xp] = o;
            }
        }
    }
}

// f *= s, elementwise over the whole padded field (the zero halo stays zero).
template <typename T>
__global__ void k_scale_const(cplx_t<T> *__restrict__ f, long long n, double s) {
#pragma clang fp contract(off)
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        cplx_t<T> v = f[e];
        v.x = (T)(s * (double)v.x);
        v.y = (T)(s * (double)v.y);
        f[e] = v;
    }
}

// odd = even * (cr + i ci) elementwise (GPUGrossPitaevskii_BCC.cu:342-343).
template <typename T>
__global__ void k_phase_copy(cplx_t<T> *__restrict__ dst, const cplx_t<T> *__restrict__ src, long long n, double cr,
                             double ci) {
#pragma clang fp contract(off)
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const cplx_t<T> v = src[e];
        const double x = (double)v.x, y = (double)v.y;
        cplx_t<T> o;
        o.x = (T)(cr * x - ci * y);
        o.y = (T)(ci * x + cr * y);
        dst[e] = o;
    }
}

}  // namespace gpe
