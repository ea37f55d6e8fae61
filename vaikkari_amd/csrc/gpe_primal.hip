// gpe_primal.hip -- Gross-Pitaevskii stepper on the BCC sites with the
// 14-point primal operator (include/vaikkari_gpe_primal.h).
//
// The Laplacian is the DEC Laplacian of 0-forms on the BCC lattice built by
// the reference's mesh library (DelaunayMesh::createBccGrid,
// jukan_koodit/Mesh/DelaunayMesh.cpp:178; Hodge weights Mesh::getEdgeHodge /
// getNodeHodge, Mesh/Mesh.cpp:1479-1489): with cube edge a, every site has 8
// diagonal neighbours (the other sublattice, weight 3/(4a^2)) and 6 axial
// neighbours (its own sublattice, weight 1/(4a^2)).  The time integrators and
// the order of the Hamiltonian's terms follow the reference stepper
// (GPUGrossPitaevskii_BCC.cu:205-215, 475-481).
//
// Field layout in HBM, the class-row SoA of the dual stepper with 2 classes:
//   site (i, j, k, s) at origin + k*sz + j*sy + s*xp + i,  sy = 2*xp,
// s = 0 the corner sublattice, s = 1 the centre sublattice.  A wave reads 64
// consecutive cells of one class row: 1 KiB (FP64) or 512 B (FP32) per
// access, and every neighbour is a class row shifted by dx + dy*sy + dz*sz.
// One zero halo cell on every side (the corner site's diagonal neighbours
// reach back one cell, the centre site's forward one cell).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/vaikkari_gpe_primal.h"
#include "gpe_hip_check.hpp"
#include "gpe_internal.hpp"
#include "gpe_fused.hpp"
#include "gpe_kernels.hpp"

using namespace gpe;

namespace {

// Trap V = 0.5 (x^2 + y^2 + rr z^2) at site positions p0 + a (i + off),
// off = 0 (corner) or 0.5 (centre); the host helpers evaluate the same
// expression, so V*dt has the same bits on both sides.
struct PTrap {
    double p0[3];
    double a;
    double rr;
};

__host__ __device__ inline double site_v(const PTrap &t, int i, int j, int k, double off) {
#pragma clang fp contract(off)
    const double x = t.p0[0] + t.a * ((double)i + off);
    const double y = t.p0[1] + t.a * ((double)j + off);
    const double z = t.p0[2] + t.a * ((double)k + off);
    return 0.5 * (x * x + y * y + t.rr * z * z);
}

// site_v split for a column of fixed (i, j): the x^2 + y^2 part once, then
// per layer the z part, in site_v's operation order (same bits).
__device__ __forceinline__ double site_xy(const PTrap &t, int i, int j, double off) {
#pragma clang fp contract(off)
    const double x = t.p0[0] + t.a * ((double)i + off);
    const double y = t.p0[1] + t.a * ((double)j + off);
    return x * x + y * y;
}
__device__ __forceinline__ double site_v_xy(const PTrap &t, double xy, int k, double off) {
#pragma clang fp contract(off)
    const double z = t.p0[2] + t.a * ((double)k + off);
    return 0.5 * (xy + t.rr * z * z);
}

// {cD, cA, c0, g} times dt in the arithmetic type.
template <typename T>
struct PCoef {
    T cD, cA, c0, g;
};

// s = dt H psi at one site from its 8 diagonal (d) and 6 axial (x)
// neighbour values.
//  FORM_REF  (FP64): cD*sum_8 d + cA*sum_6 x + ((c0 + V) + g|psi|^2) psi,
//            sums left to right from the first element.
//  FORM_DIFF (FP32): cD*sum_8 (d - psi) + cA*sum_6 (x - psi) + (V + g|psi|^2) psi,
//            algebraically equal (c0 = -(8 cD + 6 cA)) and free of the
//            cancellation between the neighbour sums and c0 psi; the products
//            are fused multiply-adds.
__device__ __forceinline__ float ffma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double ffma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T, int FORM>
__device__ __forceinline__ cplx_t<T> primal_apply(const cplx_t<T> (&d)[8], const cplx_t<T> (&x)[6], cplx_t<T> p,
                                                  T pot, const PCoef<T> &H) {
#pragma clang fp contract(off)
    cplx_t<T> s;
    const T nsq = p.x * p.x + p.y * p.y;
    T dr, di, ar, ai, c;
    if (FORM == FORM_REF) {
        dr = d[0].x;
        di = d[0].y;
#pragma unroll
        for (int q = 1; q < 8; q++) {
            dr = dr + d[q].x;
            di = di + d[q].y;
        }
        ar = x[0].x;
        ai = x[0].y;
#pragma unroll
        for (int q = 1; q < 6; q++) {
            ar = ar + x[q].x;
            ai = ai + x[q].y;
        }
        c = (H.c0 + pot) + H.g * nsq;
    } else {
        dr = d[0].x - p.x;
        di = d[0].y - p.y;
#pragma unroll
        for (int q = 1; q < 8; q++) {
            dr = dr + (d[q].x - p.x);
            di = di + (d[q].y - p.y);
        }
        ar = x[0].x - p.x;
        ai = x[0].y - p.y;
#pragma unroll
        for (int q = 1; q < 6; q++) {
            ar = ar + (x[q].x - p.x);
            ai = ai + (x[q].y - p.y);
        }
        // FP32: fused multiply-adds (one rounding fewer per product, no bit
        // target to keep)
        const T n2 = ffma(p.x, p.x, p.y * p.y);
        c = ffma(H.g, n2, pot);
        s.x = ffma(c, p.x, ffma(H.cD, dr, H.cA * ar));
        s.y = ffma(c, p.y, ffma(H.cD, di, H.cA * ai));
        return s;
    }
    s.x = (H.cD * dr + H.cA * ar) + c * p.x;
    s.y = (H.cD * di + H.cA * ai) + c * p.y;
    return s;
}

// x-neighbour exchange between adjacent lanes (XSM: 0 none, the neighbour
// is loaded; 1 ds_bpermute shuffle; 2 DPP wave shift, a VALU move with no
// memory or LDS traffic).  Lanes at a row end replace the result by a load.
enum { XS_LOAD = 0, XS_SHFL = 1, XS_DPP = 2 };
template <int XSM, typename T>
__device__ __forceinline__ cplx_t<T> from_left(cplx_t<T> v) {
    if (XSM == XS_DPP) return shift_lane_c<true>(v);
    v.x = __shfl_up(v.x, 1, 64);
    v.y = __shfl_up(v.y, 1, 64);
    return v;
}
template <int XSM, typename T>
__device__ __forceinline__ cplx_t<T> from_right(cplx_t<T> v) {
    if (XSM == XS_DPP) return shift_lane_c<false>(v);
    v.x = __shfl_down(v.x, 1, 64);
    v.y = __shfl_down(v.y, 1, 64);
    return v;
}

// Reads of the updated field for the in-place read-modify-write, issued
// together with the neighbour loads (non-temporal when NT).
template <typename T, bool NT>
__device__ __forceinline__ cplx_t<T> load_next(const cplx_t<T> *__restrict__ next, long long at) {
    cplx_t<T> n;
    if (NT) {
        n.x = __builtin_nontemporal_load(&next[at].x);
        n.y = __builtin_nontemporal_load(&next[at].y);
    } else {
        n = next[at];
    }
    return n;
}

// Finishes one site: real time next = n + (s.y, -s.x) with n the value read
// from next (GPUGrossPitaevskii_BCC.cu:215); imaginary time out = psi - s,
// streamed.  Returns |out|^2 (imaginary time) or 0.
template <typename T, int MODE, bool NT>
__device__ __forceinline__ double finish_site(cplx_t<T> *__restrict__ next, long long at, cplx_t<T> n, cplx_t<T> psi,
                                              cplx_t<T> s) {
#pragma clang fp contract(off)
    cplx_t<T> o;
    if (MODE == MODE_REAL) {
        o.x = n.x + s.y;
        o.y = n.y + (-s.x);
        if (NT) {
            __builtin_nontemporal_store(o.x, &next[at].x);
            __builtin_nontemporal_store(o.y, &next[at].y);
        } else {
            next[at] = o;
        }
        return 0.0;
    }
    o.x = psi.x - s.x;
    o.y = psi.y - s.y;
    __builtin_nontemporal_store(o.x, &next[at].x);
    __builtin_nontemporal_store(o.y, &next[at].y);
    return (double)o.x * (double)o.x + (double)o.y * (double)o.y;
}

// One cell (both sites) of an update pass; returns its sum of |out|^2
// (imaginary time).  Real time: next[site] += (s.y, -s.x) in place
// (GPUGrossPitaevskii_BCC.cu:215), non-temporal when NT.  Imaginary time:
// out = psi - s with psi = scale*prev, streamed to next.
// Neighbour order (the oracle's): diagonal dz, dy, dx over {-1, 0} (corner
// site) or {0, +1} (centre site); axial -x, +x, -y, +y, -z, +z.
// The two sites need 12 distinct class rows of prev -- corner rows (j,k),
// (j+-1,k), (j,k+-1), (j+1,k+1) and centre rows (j,k), (j-1,k), (j,k-1),
// (j-1,k-1), (j+1,k), (j,k+1) -- each loaded once, plus 10 of them shifted
// by one cell in x: from the adjacent lane (XS_SHFL, XS_DPP; the lanes at
// the two row ends load theirs, one instruction per left/right pair), or
// loaded again (XS_LOAD).  The source issues every load of the cell, the two
// reads of next included, before the first use of any of them, so a wave
// waits for memory about once per cell rather than once per dependent step
// (with XS_LOAD's 24 loads the compiler interleaves some waits to save
// registers).
template <typename T, int FORM, int MODE, int TY, bool NT, int XSM>
__device__ __forceinline__ double primal_cell(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                              const Layout &L, const PCoef<T> &H, const PTrap &trap, double dt,
                                              const double *__restrict__ scale, int bx, int by, int bz, int lsh) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    const int lw = 1 << lsh;
    const int xl = threadIdx.x & (lw - 1);
    const int i = (bx << lsh) + xl;
    const int j = ((by * TY + threadIdx.y) << (6 - lsh)) + (threadIdx.x >> lsh);
    const int k = bz;
    const bool act = i < L.nx && j < L.ny;
    // a lane at a row end takes its x-neighbour from memory (the halo at the grid's edge)
    const bool left_end = act && xl == 0, right_end = act && xl == lw - 1;
    const long long sy = L.sy, sz = L.sz;
    const long long b0 = L.origin + (long long)k * sz + (long long)j * sy + i;  // corner site
    const long long b1 = b0 + L.xp;                                               // centre site
    T sc = (T)1;
    if (MODE == MODE_IMAG) sc = (T)(*scale);
    auto ld = [&](long long at) {
        C v = prev[at];
        if (MODE == MODE_IMAG) v = cscale<T>(v, sc);
        return v;
    };
    C zero;
    zero.x = (T)0;
    zero.y = (T)0;
    // lanes beyond the grid hold zeros: the halo value a neighbouring lane needs
    C c0 = zero, c1 = zero, cym = zero, cyp = zero, czm = zero, czp = zero, u0 = zero, u1 = zero, u2 = zero;
    C v3 = zero, zyp = zero, zzp = zero;
    C w0, w1, w2, w3, xm, xq, e1, e2, e3, zxp;  // x-shifted rows: w* centre rows at i-1, e* corner rows at i+1
    C t[5], n0, n1;
    if (act) {
        c0 = ld(b0);
        c1 = ld(b1);
        u0 = ld(b1 - sz - sy);  // centre rows (j-1,k-1), (j,k-1), (j-1,k)
        u1 = ld(b1 - sz);
        u2 = ld(b1 - sy);
        cym = ld(b0 - sy);  // corner rows (j-1,k), (j+1,k), (j,k-1), (j,k+1)
        cyp = ld(b0 + sy);
        czm = ld(b0 - sz);
        czp = ld(b0 + sz);
        v3 = ld(b0 + sz + sy);  // corner row (j+1,k+1)
        zyp = ld(b1 + sy);      // centre rows (j+1,k), (j,k+1)
        zzp = ld(b1 + sz);
        if (XSM == XS_LOAD) {
            w0 = ld(b1 - sz - sy - 1);
            w1 = ld(b1 - sz - 1);
            w2 = ld(b1 - sy - 1);
            w3 = ld(b1 - 1);
            xm = ld(b0 - 1);
            xq = ld(b0 + 1);
            e1 = ld(b0 + sy + 1);
            e2 = ld(b0 + sz + 1);
            e3 = ld(b0 + sz + sy + 1);
            zxp = ld(b1 + 1);
        }
    }
    if (XSM != XS_LOAD && (left_end || right_end)) {
        t[0] = ld(left_end ? b1 - sz - sy - 1 : b0 + 1);  // w0 | xq
        t[1] = ld(left_end ? b1 - sz - 1 : b0 + sy + 1);  // w1 | e1
        t[2] = ld(left_end ? b1 - sy - 1 : b0 + sz + 1);  // w2 | e2
        t[3] = ld(left_end ? b1 - 1 : b0 + sz + sy + 1);  // w3 | e3
        t[4] = ld(left_end ? b0 - 1 : b1 + 1);            // xm | zxp
    }
    if (MODE == MODE_REAL && act) {
        n0 = load_next<T, NT>(next, b0);
        n1 = load_next<T, NT>(next, b1);
    }
    if (XSM != XS_LOAD) {
        w0 = from_left<XSM, T>(u0);
        w1 = from_left<XSM, T>(u1);
        w2 = from_left<XSM, T>(u2);
        w3 = from_left<XSM, T>(c1);
        xm = from_left<XSM, T>(c0);
        xq = from_right<XSM, T>(c0);
        e1 = from_right<XSM, T>(cyp);
        e2 = from_right<XSM, T>(czp);
        e3 = from_right<XSM, T>(v3);
        zxp = from_right<XSM, T>(c1);
        if (left_end) {
            w0 = t[0];
            w1 = t[1];
            w2 = t[2];
            w3 = t[3];
            xm = t[4];
        }
        if (right_end) {
            xq = t[0];
            e1 = t[1];
            e2 = t[2];
            e3 = t[3];
            zxp = t[4];
        }
    }
    double part = 0.0;
    if (act) {
        const C dC[8] = {w0, u0, w1, u1, w2, u2, w3, c1};
        const C aC[6] = {xm, xq, cym, cyp, czm, czp};
        const C s0 = primal_apply<T, FORM>(dC, aC, c0, (T)(site_v(trap, i, j, k, 0.0) * dt), H);
        const C dZ[8] = {c0, xq, cyp, e1, czp, e2, v3, e3};
        const C aZ[6] = {w3, zxp, u2, zyp, u1, zzp};
        const C s1 = primal_apply<T, FORM>(dZ, aZ, c1, (T)(site_v(trap, i, j, k, 0.5) * dt), H);
        part = finish_site<T, MODE, NT>(next, b0, n0, c0, s0);
        part += finish_site<T, MODE, NT>(next, b1, n1, c1, s1);
    }
    return part;
}

// Update pass: one thread per cell, block (64, TY), logical grid
// (gx, gy, nz) ordered over the XCDs by block_coords.
template <typename T, int FORM, int MODE, int TY, bool NT, int XSM>
__global__ void __launch_bounds__(64 * TY) k_primal(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                                    Layout L, PCoef<T> H, PTrap trap, double dt,
                                                    const double *__restrict__ scale, double *__restrict__ partials,
                                                    Grid3 G) {
    __shared__ double lds[TY];
    int bx, by, bz;
    block_coords(G, bx, by, bz);
    double part[1];
    part[0] = primal_cell<T, FORM, MODE, TY, NT, XSM>(next, prev, L, H, trap, dt, scale, bx, by, bz, G.lsh);
    if (MODE == MODE_IMAG) {
        // partials in logical tile order (independent of the block-to-XCD order)
        block_sum<1, TY>(part, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[bx + G.gx * (by + G.gy * bz)] = part[0];
    }
}

// Z-MARCH: one thread per (i, j) column walking CZ layers upwards.  Of the
// 12 class rows a cell needs, the previous layer's rows stay in registers
// (own corner and centre, the centre rows (j, k-1) and (j-1, k-1), the corner
// row (j+1, k) fetched as (j+1, k+1) one layer earlier), and so do 4 of the
// 10 x-shifted copies; per layer a lane issues 6 row loads -- the new layer's
// corner rows (j, k+1), (j+1, k+1) and centre row (j, k+1), and the
// y-neighbours corner (j-1, k), centre (j-+1, k), which the neighbouring
// waves of the block read too -- and takes the 6 new x-shifts by DPP.  The
// lanes at the two row ends fetch their x-neighbours with 3 shared loads
// (left end: x-1 values, right end: x+1 values, one instruction per pair).
// All loads of layer k+1, the reads of next included, are issued before
// layer k is computed (one layer of software prefetch), so memory latency
// overlaps the arithmetic.  CZ is small (4 to 16) because the layer loop is
// fully unrolled.  Same arithmetic and neighbour order as
// primal_cell: bit-identical.
template <typename T>
struct ZLayer {
    cplx_t<T> czp, v3, zzp, cym, u2, zyp;  // rows of prev
    cplx_t<T> t0, t1, t2;                  // row-end x-neighbours
    cplx_t<T> n0, n1;                      // next, for the read-modify-write
};

template <typename T, int FORM, int MODE, int TY, int CZ, bool NT, bool LY>
__device__ __forceinline__ double primal_column(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                                const Layout &L, const PCoef<T> &H, const PTrap &trap, double dt,
                                                const double *__restrict__ scale, int bx, int by, int bz, int lsh) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    const int lw = 1 << lsh;
    const int xl = threadIdx.x & (lw - 1);
    const int i = (bx << lsh) + xl;
    const int j = ((by * TY + threadIdx.y) << (6 - lsh)) + (threadIdx.x >> lsh);
    const int k0 = bz * CZ;
    const int k1 = min(k0 + CZ, L.nz);
    const bool act = i < L.nx && j < L.ny;
    const bool left_end = act && xl == 0, right_end = act && xl == lw - 1;
    const long long xp = L.xp, sy = L.sy, sz = L.sz;
    // Loads are unconditional, so the layer code is straight-line and the
    // compiler's wait counts stay exact: a lane beyond the grid reads the
    // zero halo column x = -1 of an existing row instead (its row loads then
    // give the zeros its neighbours' x-shifts need), and the lanes not at a
    // row end point their row-end load at their own corner site.
    const int ie = act ? i : -1;
    const int je = act ? j : min(j, L.ny - 1);
    T sc = (T)1;
    if (MODE == MODE_IMAG) sc = (T)(*scale);
    auto ld = [&](long long at) {
        C v = prev[at];
        if (MODE == MODE_IMAG) v = cscale<T>(v, sc);
        return v;
    };
    // loads of one layer (b0: corner site of (ie, je, k))
    auto fetch = [&](long long b0, ZLayer<T> &f) {
        const long long b1 = b0 + xp;
        f.czp = ld(b0 + sz);      // corner (j, k+1)
        f.v3 = ld(b0 + sz + sy);  // corner (j+1, k+1)
        f.zzp = ld(b1 + sz);      // centre (j, k+1)
        if (!LY) {
            f.cym = ld(b0 - sy);  // corner (j-1, k)
            f.u2 = ld(b1 - sy);   // centre (j-1, k)
            f.zyp = ld(b1 + sy);  // centre (j+1, k)
        }
        f.t0 = ld(left_end ? b1 - sy - 1 : right_end ? b0 + sz + 1 : b0);       // w2 | e2
        f.t1 = ld(left_end ? b1 - 1 : right_end ? b0 + sz + sy + 1 : b0);       // w3 | e3
        f.t2 = ld(left_end ? b0 - 1 : right_end ? b1 + 1 : b0);                 // xm | zxp
        if (MODE == MODE_REAL) {
            f.n0 = load_next<T, NT>(next, b0);
            f.n1 = load_next<T, NT>(next, b1);
        }
    };
    long long b0 = L.origin + (long long)k0 * sz + (long long)je * sy + ie;
    // carried state at layer k
    C c0 = ld(b0), c1 = ld(b0 + xp), czm = ld(b0 - sz), u1 = ld(b0 + xp - sz), u0 = ld(b0 + xp - sz - sy);
    C cyp = ld(b0 + sy);
    const C t0 = ld(left_end ? b0 + xp - sz - sy - 1 : right_end ? b0 + sy + 1 : b0);  // w0 | e1
    const C t1 = ld(left_end ? b0 + xp - sz - 1 : right_end ? b0 + 1 : b0);            // w1 | xq
    // two layer buffers used in turn (the loop is unrolled, so buf[kk & 1]
    // names fixed registers and no copy waits for a load in flight)
    ZLayer<T> buf[2];
    fetch(b0, buf[0]);
    C w0 = shift_lane_c<true>(u0), w1 = shift_lane_c<true>(u1);
    C e1 = shift_lane_c<false>(cyp), xq = shift_lane_c<false>(c0);
    if (left_end) {
        w0 = t0;
        w1 = t1;
    }
    if (right_end) {
        e1 = t0;
        xq = t1;
    }
    const double xy0 = site_xy(trap, i, j, 0.0), xy1 = site_xy(trap, i, j, 0.5);
    double part = 0.0;
    // unrolled: with no loop back edge the compiler's wait counts track each
    // load, so the wait for layer k leaves the prefetch of layer k+1 in flight
#pragma unroll
    for (int kk = 0; kk < CZ; kk++, b0 += sz) {
        const int k = k0 + kk;
        if (k >= k1) break;
        ZLayer<T> &cur = buf[kk & 1];
        if (LY) {
            // the y-neighbour rows of layer k (rows the block's other waves
            // read too) are loaded in the layer itself, before the prefetch so
            // that waiting for them leaves the prefetch in flight: three fewer
            // values held across a layer
            cur.cym = ld(b0 - sy);
            cur.u2 = ld(b0 + xp - sy);
            cur.zyp = ld(b0 + xp + sy);
        }
        if (k + 1 < k1) fetch(b0 + sz, buf[(kk + 1) & 1]);
        // keep the scheduler from hoisting layer k's first uses (and their
        // wait) above the prefetch loads
        __builtin_amdgcn_sched_barrier(0);
        C w2 = shift_lane_c<true>(cur.u2), w3 = shift_lane_c<true>(c1), xm = shift_lane_c<true>(c0);
        C e2 = shift_lane_c<false>(cur.czp), e3 = shift_lane_c<false>(cur.v3), zxp = shift_lane_c<false>(c1);
        if (left_end) {
            w2 = cur.t0;
            w3 = cur.t1;
            xm = cur.t2;
        }
        if (right_end) {
            e2 = cur.t0;
            e3 = cur.t1;
            zxp = cur.t2;
        }
        const C dC[8] = {w0, u0, w1, u1, w2, cur.u2, w3, c1};
        const C aC[6] = {xm, xq, cur.cym, cyp, czm, cur.czp};
        const C s0 = primal_apply<T, FORM>(dC, aC, c0, (T)(site_v_xy(trap, xy0, k, 0.0) * dt), H);
        const C dZ[8] = {c0, xq, cyp, e1, cur.czp, e2, cur.v3, e3};
        const C aZ[6] = {w3, zxp, cur.u2, cur.zyp, u1, cur.zzp};
        const C s1 = primal_apply<T, FORM>(dZ, aZ, c1, (T)(site_v_xy(trap, xy1, k, 0.5) * dt), H);
        if (act) {
            part += finish_site<T, MODE, NT>(next, b0, cur.n0, c0, s0);
            part += finish_site<T, MODE, NT>(next, b0 + xp, cur.n1, c1, s1);
        }
        // layer k+1 from what layer k read
        czm = c0;
        c0 = cur.czp;
        u1 = c1;
        c1 = cur.zzp;
        u0 = cur.u2;
        w0 = w2;
        w1 = w3;
        cyp = cur.v3;
        e1 = e3;
        xq = e2;
    }
    return part;
}

template <typename T, int FORM, int MODE, int TY, int CZ, bool NT, bool LY>
__global__ void __launch_bounds__(64 * TY) k_primal_zm(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                                       Layout L, PCoef<T> H, PTrap trap, double dt,
                                                       const double *__restrict__ scale, double *__restrict__ partials,
                                                       Grid3 G) {
    __shared__ double lds[TY];
    int bx, by, bz;
    block_coords(G, bx, by, bz);
    double part[1];
    part[0] = primal_column<T, FORM, MODE, TY, CZ, NT, LY>(next, prev, L, H, trap, dt, scale, bx, by, bz, G.lsh);
    if (MODE == MODE_IMAG) {
        block_sum<1, TY>(part, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[bx + G.gx * (by + G.gy * bz)] = part[0];
    }
}

// Z-MARCH with buffer-resource addressing: the same column walk, layer
// order, prefetch and arithmetic as primal_column, but every access is a raw
// buffer load or store with one per-lane 32-bit byte offset (the column) and a
// wave-uniform one (layer and row, advanced by scalar adds), so a row access
// costs no 64-bit vector address arithmetic.  Fields must be smaller than
// 2 GiB (the host checks).  Lanes without a cell store to an offset beyond the
// buffer, which the buffer's range check discards, so stores need no branch.
// FP32 (FORM_DIFF) takes V*dt as (0.5 (x^2 + y^2) dt) of the column plus
// (0.5 rr z^2 dt) of the layer from a per-layer table, both rounded to FP32
// (about 1 ulp from the FP64 expression); FP64 evaluates site_v exactly as
// primal_column, so its results keep the same bits.
template <typename T>
struct BufC;
template <>
struct BufC<float> {
    template <int AUX>
    static __device__ __forceinline__ cplx_t<float> ld(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(cplx_t<float>, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, AUX));
    }
    template <int AUX>
    static __device__ __forceinline__ void st(cplx_t<float> x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), x), r, v, s, AUX);
    }
};
template <>
struct BufC<double> {
    template <int AUX>
    static __device__ __forceinline__ cplx_t<double> ld(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(cplx_t<double>, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, AUX));
    }
    template <int AUX>
    static __device__ __forceinline__ void st(cplx_t<double> x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        buffer_store_128<AUX>(x, r, v, s);
    }
};
constexpr int kBufNT = 2;             // cache-policy bits of a non-temporal buffer access
constexpr int kBufWord3 = 0x00020000;  // raw buffer, 32-bit data format
constexpr uint32_t kBufDiscard = 0x80000000u;  // byte offset beyond every buffer this kernel addresses

template <typename T, int FORM, int MODE, int TY, int CZ, bool NT, bool LY, int PD>
__device__ __forceinline__ double primal_column_b(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev,
                                                  const Layout &L, const PCoef<T> &H, const PTrap &trap, double dt,
                                                  const float *__restrict__ vzt, const double *__restrict__ scale,
                                                  int bx, int by, int bz, int lsh, int czr) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    constexpr uint32_t ES = sizeof(C);
    const int bytes = (int)((long long)L.elems * ES);
    const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc((void *)prev, (short)0, bytes, kBufWord3);
    const __amdgpu_buffer_rsrc_t rn = __builtin_amdgcn_make_buffer_rsrc((void *)next, (short)0, bytes, kBufWord3);
    const int lw = 1 << lsh;
    const int xl = threadIdx.x & (lw - 1);
    const int i = (bx << lsh) + xl;
    const int j = ((by * TY + threadIdx.y) << (6 - lsh)) + (threadIdx.x >> lsh);
    // CZ > 0: layers per chunk at compile time; CZ = 0: czr, from the launch
    const int cz = CZ > 0 ? CZ : czr;
    const int k0 = bz * cz;
    const int k1 = min(k0 + cz, L.nz);
    const bool act = i < L.nx && j < L.ny;
    const bool left_end = act && xl == 0, right_end = act && xl == lw - 1;
    const uint32_t xp = (uint32_t)L.xp, sy = (uint32_t)L.sy, sz = (uint32_t)L.sz;
    const int ie = act ? i : -1;
    const int je = act ? j : min(j, L.ny - 1);
    // element of row (dz, dy, s) of layer k: cb + k sz + R(dz, dy, s), every term >= 0
    const uint32_t cb = (uint32_t)(L.origin - L.sz - L.sy + (long long)je * L.sy + ie);
    auto R = [&](int dz, int dy, int sl) -> uint32_t {
        return (uint32_t)(dz + 1) * sz + (uint32_t)(dy + 1) * sy + (uint32_t)sl * xp;
    };
    // the same row offsets as wave-uniform values: the compiler may share an
    // R(...) with the per-lane offsets below, and a row offset it holds in a
    // vector register would turn each load into a loop over distinct values
    auto RS = [&](int dz, int dy, int sl) -> uint32_t { return __builtin_amdgcn_readfirstlane(R(dz, dy, sl)); };
    const uint32_t vcol = cb * ES;                              // per-lane byte offset
    const uint32_t vst = act ? vcol : kBufDiscard;              // stores of lanes without a cell are discarded
    // row-end loads: per-lane byte offsets (left end: x-1 values, right end:
    // x+1 values); the other lanes take an offset beyond the buffer, whose
    // range check returns zero without a cache access, so these loads cost
    // the memory pipeline two lanes' lines instead of a whole row's
    auto vend = [&](uint32_t l, uint32_t r) -> uint32_t {
        return left_end ? (cb + l - 1) * ES : right_end ? (cb + r + 1) * ES : kBufDiscard;
    };
    const uint32_t vt0 = vend(R(0, -1, 1), R(1, 0, 0));
    const uint32_t vt1 = vend(R(0, 0, 1), R(1, 1, 0));
    const uint32_t vt2 = vend(R(0, 0, 0), R(0, 0, 1));
    const uint32_t vi0 = vend(R(-1, -1, 1), R(0, 1, 0));
    const uint32_t vi1 = vend(R(-1, 0, 1), R(0, 0, 0));
    T sc = (T)1;
    if (MODE == MODE_IMAG) sc = (T)(*scale);
    auto ldv = [&](uint32_t v, uint32_t soff) {
        C x = BufC<T>::template ld<0>(rp, v, soff);
        if (MODE == MODE_IMAG) x = cscale<T>(x, sc);
        return x;
    };
    auto ld = [&](uint32_t lay, int dz, int dy, int sl) { return ldv(vcol, (lay + RS(dz, dy, sl)) * ES); };
    auto fetch = [&](uint32_t lay, ZLayer<T> &f) {  // lay = k sz of the layer
        f.czp = ld(lay, 1, 0, 0);
        f.v3 = ld(lay, 1, 1, 0);
        f.zzp = ld(lay, 1, 0, 1);
        if (!LY) {
            f.cym = ld(lay, 0, -1, 0);
            f.u2 = ld(lay, 0, -1, 1);
            f.zyp = ld(lay, 0, 1, 1);
        }
        f.t0 = ldv(vt0, lay * ES);
        f.t1 = ldv(vt1, lay * ES);
        f.t2 = ldv(vt2, lay * ES);
        if (MODE == MODE_REAL) {
            f.n0 = BufC<T>::template ld<NT ? kBufNT : 0>(rn, vcol, (lay + RS(0, 0, 0)) * ES);
            f.n1 = BufC<T>::template ld<NT ? kBufNT : 0>(rn, vcol, (lay + RS(0, 0, 1)) * ES);
        }
    };
    uint32_t lay = (uint32_t)k0 * sz;
    C c0 = ld(lay, 0, 0, 0), c1 = ld(lay, 0, 0, 1), czm = ld(lay, -1, 0, 0), u1 = ld(lay, -1, 0, 1),
      u0 = ld(lay, -1, -1, 1);
    C cyp = ld(lay, 0, 1, 0);
    const C t0 = ldv(vi0, lay * ES);
    const C t1 = ldv(vi1, lay * ES);
    // PD + 1 layer buffers used in turn: the loads of layer k + PD are issued
    // while layer k is computed (the loop is unrolled, so buf[kk % (PD + 1)]
    // names fixed registers)
    ZLayer<T> buf[PD + 1];
#pragma unroll
    for (int q = 0; q < PD; q++) fetch(lay + (uint32_t)q * sz, buf[q]);
    C w0 = shift_lane_c<true>(u0), w1 = shift_lane_c<true>(u1);
    C e1 = shift_lane_c<false>(cyp), xq = shift_lane_c<false>(c0);
    if (left_end) {
        w0 = t0;
        w1 = t1;
    }
    if (right_end) {
        e1 = t0;
        xq = t1;
    }
    const double xy0 = site_xy(trap, i, j, 0.0), xy1 = site_xy(trap, i, j, 0.5);
    const T vxy0 = (T)(0.5 * xy0 * dt), vxy1 = (T)(0.5 * xy1 * dt);
    double part = 0.0;
    // layer k (lay = k sz) from the buffer cur, prefetching layer k + PD into nxt
    auto step = [&](int k, uint32_t lay, ZLayer<T> &cur, ZLayer<T> &nxt) {
        if (LY) {
            cur.cym = ld(lay, 0, -1, 0);
            cur.u2 = ld(lay, 0, -1, 1);
            cur.zyp = ld(lay, 0, 1, 1);
        }
        // the prefetch also runs at the chunk's last layers (a branch here would
        // make the compiler merge the buffers with register copies): beyond the
        // halo layer nz + 1 the buffer's range check returns zeros, never used
        fetch(lay + (uint32_t)PD * sz, nxt);
        __builtin_amdgcn_sched_barrier(0);
        C w2 = shift_lane_c<true>(cur.u2), w3 = shift_lane_c<true>(c1), xm = shift_lane_c<true>(c0);
        C e2 = shift_lane_c<false>(cur.czp), e3 = shift_lane_c<false>(cur.v3), zxp = shift_lane_c<false>(c1);
        if (left_end) {
            w2 = cur.t0;
            w3 = cur.t1;
            xm = cur.t2;
        }
        if (right_end) {
            e2 = cur.t0;
            e3 = cur.t1;
            zxp = cur.t2;
        }
        T v0, v1;
        if (FORM == FORM_REF) {
            v0 = (T)(site_v_xy(trap, xy0, k, 0.0) * dt);
            v1 = (T)(site_v_xy(trap, xy1, k, 0.5) * dt);
        } else {
            v0 = vxy0 + (T)vzt[2 * k];
            v1 = vxy1 + (T)vzt[2 * k + 1];
        }
        const C dC[8] = {w0, u0, w1, u1, w2, cur.u2, w3, c1};
        const C aC[6] = {xm, xq, cur.cym, cyp, czm, cur.czp};
        const C s0 = primal_apply<T, FORM>(dC, aC, c0, v0, H);
        const C dZ[8] = {c0, xq, cyp, e1, cur.czp, e2, cur.v3, e3};
        const C aZ[6] = {w3, zxp, cur.u2, cur.zyp, u1, cur.zzp};
        const C s1 = primal_apply<T, FORM>(dZ, aZ, c1, v1, H);
        C o0, o1;
        if (MODE == MODE_REAL) {
            o0.x = cur.n0.x + s0.y;
            o0.y = cur.n0.y + (-s0.x);
            o1.x = cur.n1.x + s1.y;
            o1.y = cur.n1.y + (-s1.x);
            BufC<T>::template st<NT ? kBufNT : 0>(o0, rn, vst, (lay + RS(0, 0, 0)) * ES);
            BufC<T>::template st<NT ? kBufNT : 0>(o1, rn, vst, (lay + RS(0, 0, 1)) * ES);
        } else {
            o0.x = c0.x - s0.x;
            o0.y = c0.y - s0.y;
            o1.x = c1.x - s1.x;
            o1.y = c1.y - s1.y;
            BufC<T>::template st<kBufNT>(o0, rn, vst, (lay + RS(0, 0, 0)) * ES);
            BufC<T>::template st<kBufNT>(o1, rn, vst, (lay + RS(0, 0, 1)) * ES);
            if (act)
                part += ((double)o0.x * (double)o0.x + (double)o0.y * (double)o0.y) +
                        ((double)o1.x * (double)o1.x + (double)o1.y * (double)o1.y);
        }
        czm = c0;
        c0 = cur.czp;
        u1 = c1;
        c1 = cur.zzp;
        u0 = cur.u2;
        w0 = w2;
        w1 = w3;
        cyp = cur.v3;
        e1 = e3;
        xq = e2;
    };
    if constexpr (CZ > 0) {
#pragma unroll
        for (int kk = 0; kk < CZ; kk++, lay += sz) {
            if (k0 + kk >= k1) break;
            step(k0 + kk, lay, buf[kk % (PD + 1)], buf[(kk + PD) % (PD + 1)]);
        }
    } else {
        // a run-time chunk: groups of PD + 1 layers, so that every buffer
        // keeps fixed registers, then the rest
        int k = k0;
        for (; k + PD < k1; k += PD + 1, lay += (uint32_t)(PD + 1) * sz) {
#pragma unroll
            for (int q = 0; q <= PD; q++) step(k + q, lay + (uint32_t)q * sz, buf[q], buf[(q + PD) % (PD + 1)]);
        }
#pragma unroll
        for (int q = 0; q < PD; q++)
            if (k + q < k1) step(k + q, lay + (uint32_t)q * sz, buf[q], buf[(q + PD) % (PD + 1)]);
    }
    return part;
}

template <typename T, int FORM, int MODE, int TY, int CZ, bool NT, bool LY, int PD>
__global__ void __launch_bounds__(64 * TY)
    k_primal_zb(cplx_t<T> *__restrict__ next, const cplx_t<T> *__restrict__ prev, Layout L, PCoef<T> H, PTrap trap,
                double dt, const float *__restrict__ vzt, const double *__restrict__ scale,
                double *__restrict__ partials, Grid3 G) {
    __shared__ double lds[TY];
    int bx, by, bz;
    block_coords(G, bx, by, bz);
    double part[1];
    part[0] = primal_column_b<T, FORM, MODE, TY, CZ, NT, LY, PD>(next, prev, L, H, trap, dt, vzt, scale, bx, by, bz,
                                                                 G.lsh, G.cz);
    if (MODE == MODE_IMAG) {
        block_sum<1, TY>(part, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[bx + G.gx * (by + G.gy * bz)] = part[0];
    }
}

// Z-MARCH, FP32, two cells per lane: as primal_column_b (buffer addressing,
// one layer of prefetch, the same neighbour order and arithmetic per site),
// but lane l of a wave holds the cells 2l and 2l + 1 of a 128-cell row, so
// every row access is one 16-byte load or store per lane: half the
// vector-memory instructions, and half the texture-address work, of one cell
// per lane (the FP32 z-march is bound by that unit: TA busy about 90 % of
// the kernel time at 256^3).  x-neighbours inside the lane are the other
// cell, across lanes one DPP wave shift; the row-end lanes load the cell
// beyond their end.  A lane whose second cell is the zero halo column x = nx
// stores zero there; a lane whose first cell is beyond the grid discards
// its stores.  Wave rows always span 128 cells (lsh = 6).
template <typename T>
struct CP {
    cplx_t<T> a, b;  // cells 2l, 2l + 1
};
__device__ __forceinline__ CP<float> ld_pair(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
    return __builtin_bit_cast(CP<float>, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
}
template <int AUX>
__device__ __forceinline__ CP<float> ld_pair_aux(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
    return __builtin_bit_cast(CP<float>, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, AUX));
}
template <int AUX>
__device__ __forceinline__ void st_pair(CP<float> x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
    buffer_store_128<AUX>(x, r, v, s);
}
// x-neighbours of both cells: from the left (x - 1) and from the right (x + 1)
__device__ __forceinline__ CP<float> pair_from_left(const CP<float> &v) {
    CP<float> o;
    o.a = shift_lane_c<true>(v.b);
    o.b = v.a;
    return o;
}
__device__ __forceinline__ CP<float> pair_from_right(const CP<float> &v) {
    CP<float> o;
    o.a = v.b;
    o.b = shift_lane_c<false>(v.a);
    return o;
}

struct ZLayer2 {
    CP<float> czp, v3, zzp, cym, u2, zyp;  // rows of prev
    cplx_t<float> t0, t1, t2;              // row-end x-neighbours (one cell)
    CP<float> n0, n1;                      // next, for the read-modify-write
};

// LY: 0 every load of a layer is prefetched a layer ahead; 1 the y-neighbour
// rows are loaded in the layer itself (fewer registers); 2 the row-end loads
// as well.
template <int MODE, int TY, int CZ, bool NT, int LY>
__device__ __forceinline__ double primal_column_p(cplx_t<float> *__restrict__ next,
                                                  const cplx_t<float> *__restrict__ prev, const Layout &L,
                                                  const PCoef<float> &H, const PTrap &trap, double dt,
                                                  const float *__restrict__ vzt, const double *__restrict__ scale,
                                                  int bx, int by, int bz, int czr) {
#pragma clang fp contract(off)
    using C = cplx_t<float>;
    using P = CP<float>;
    constexpr uint32_t ES = sizeof(C);
    const int bytes = (int)((long long)L.elems * ES);
    const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc((void *)prev, (short)0, bytes, kBufWord3);
    const __amdgpu_buffer_rsrc_t rn = __builtin_amdgcn_make_buffer_rsrc((void *)next, (short)0, bytes, kBufWord3);
    const int lane = threadIdx.x;
    const int i = bx * 128 + 2 * lane;  // first cell of the lane
    const int j = by * TY + threadIdx.y;
    const int cz = CZ > 0 ? CZ : czr;  // CZ = 0: layers per chunk from the launch
    const int k0 = bz * cz;
    const int k1 = min(k0 + cz, L.nz);
    const bool row = j < L.ny;
    const bool actA = row && i < L.nx, actB = row && i + 1 < L.nx;
    const bool left_end = row && lane == 0, right_end = row && lane == 63;
    const uint32_t xp = (uint32_t)L.xp, sy = (uint32_t)L.sy, sz = (uint32_t)L.sz;
    const int je = row ? j : L.ny - 1;
    // lanes past the row end read whatever follows it (inside the buffer, or
    // zeros past its end): those values reach only cells beyond the grid
    const uint32_t cb = (uint32_t)(L.origin - L.sz - L.sy + (long long)je * L.sy + i);
    auto R = [&](int dz, int dy, int sl) -> uint32_t {
        return (uint32_t)(dz + 1) * sz + (uint32_t)(dy + 1) * sy + (uint32_t)sl * xp;
    };
    auto RS = [&](int dz, int dy, int sl) -> uint32_t { return __builtin_amdgcn_readfirstlane(R(dz, dy, sl)); };
    const uint32_t vcol = cb * ES;
    const uint32_t vst = actA ? vcol : kBufDiscard;
    // row-end loads (one cell): left end x - 1 of its first cell, right end
    // x + 1 of its second; the other lanes an offset beyond the buffer (zero,
    // no cache access)
    auto vend = [&](uint32_t l, uint32_t r) -> uint32_t {
        return left_end ? (cb + l - 1) * ES : right_end ? (cb + r + 2) * ES : kBufDiscard;
    };
    const uint32_t vt0 = vend(R(0, -1, 1), R(1, 0, 0));
    const uint32_t vt1 = vend(R(0, 0, 1), R(1, 1, 0));
    const uint32_t vt2 = vend(R(0, 0, 0), R(0, 0, 1));
    const uint32_t vi0 = vend(R(-1, -1, 1), R(0, 1, 0));
    const uint32_t vi1 = vend(R(-1, 0, 1), R(0, 0, 0));
    float sc = 1.0f;
    if (MODE == MODE_IMAG) sc = (float)(*scale);
    auto scl = [&](P v) {
        if (MODE == MODE_IMAG) {
            v.a = cscale<float>(v.a, sc);
            v.b = cscale<float>(v.b, sc);
        }
        return v;
    };
    auto ld = [&](uint32_t lay, int dz, int dy, int sl) { return scl(ld_pair(rp, vcol, (lay + RS(dz, dy, sl)) * ES)); };
    auto ld1 = [&](uint32_t v, uint32_t lay) {
        C x = BufC<float>::template ld<0>(rp, v, lay * ES);
        if (MODE == MODE_IMAG) x = cscale<float>(x, sc);
        return x;
    };
    auto fetch = [&](uint32_t lay, ZLayer2 &f) {
        f.czp = ld(lay, 1, 0, 0);
        f.v3 = ld(lay, 1, 1, 0);
        f.zzp = ld(lay, 1, 0, 1);
        if (!LY) {
            f.cym = ld(lay, 0, -1, 0);
            f.u2 = ld(lay, 0, -1, 1);
            f.zyp = ld(lay, 0, 1, 1);
        }
        if (LY < 2) {
            f.t0 = ld1(vt0, lay);
            f.t1 = ld1(vt1, lay);
            f.t2 = ld1(vt2, lay);
        }
        if (MODE == MODE_REAL) {
            f.n0 = ld_pair_aux<NT ? kBufNT : 0>(rn, vcol, (lay + RS(0, 0, 0)) * ES);
            f.n1 = ld_pair_aux<NT ? kBufNT : 0>(rn, vcol, (lay + RS(0, 0, 1)) * ES);
        }
    };
    uint32_t lay = (uint32_t)k0 * sz;
    P c0 = ld(lay, 0, 0, 0), c1 = ld(lay, 0, 0, 1), czm = ld(lay, -1, 0, 0), u1 = ld(lay, -1, 0, 1),
      u0 = ld(lay, -1, -1, 1);
    P cyp = ld(lay, 0, 1, 0);
    const C t0 = ld1(vi0, lay);
    const C t1 = ld1(vi1, lay);
    ZLayer2 buf[2];
    fetch(lay, buf[0]);
    P w0 = pair_from_left(u0), w1 = pair_from_left(u1);
    P e1 = pair_from_right(cyp), xq = pair_from_right(c0);
    if (left_end) {
        w0.a = t0;
        w1.a = t1;
    }
    if (right_end) {
        e1.b = t0;
        xq.b = t1;
    }
    const double yy = trap.p0[1] + trap.a * (double)j, yy1 = trap.p0[1] + trap.a * ((double)j + 0.5);
    const double xa = trap.p0[0] + trap.a * (double)i, xb = trap.p0[0] + trap.a * (double)(i + 1);
    const double xa1 = trap.p0[0] + trap.a * ((double)i + 0.5), xb1 = trap.p0[0] + trap.a * ((double)(i + 1) + 0.5);
    const float vA0 = (float)(0.5 * (xa * xa + yy * yy) * dt), vB0 = (float)(0.5 * (xb * xb + yy * yy) * dt);
    const float vA1 = (float)(0.5 * (xa1 * xa1 + yy1 * yy1) * dt), vB1 = (float)(0.5 * (xb1 * xb1 + yy1 * yy1) * dt);
    double part = 0.0;
    // layer k (lay = k sz) from the buffer cur, prefetching layer k + 1 into nxt
    auto step = [&](int k, uint32_t lay, ZLayer2 &cur, ZLayer2 &nxt) {
        if (LY) {
            cur.cym = ld(lay, 0, -1, 0);
            cur.u2 = ld(lay, 0, -1, 1);
            cur.zyp = ld(lay, 0, 1, 1);
        }
        if (LY == 2) {
            cur.t0 = ld1(vt0, lay);
            cur.t1 = ld1(vt1, lay);
            cur.t2 = ld1(vt2, lay);
        }
        fetch(lay + sz, nxt);
        __builtin_amdgcn_sched_barrier(0);
        P w2 = pair_from_left(cur.u2), w3 = pair_from_left(c1), xm = pair_from_left(c0);
        P e2 = pair_from_right(cur.czp), e3 = pair_from_right(cur.v3), zxp = pair_from_right(c1);
        if (left_end) {
            w2.a = cur.t0;
            w3.a = cur.t1;
            xm.a = cur.t2;
        }
        if (right_end) {
            e2.b = cur.t0;
            e3.b = cur.t1;
            zxp.b = cur.t2;
        }
        const float vz0 = vzt[2 * k], vz1 = vzt[2 * k + 1];
        P o0, o1;
        {
            const C dC[8] = {w0.a, u0.a, w1.a, u1.a, w2.a, cur.u2.a, w3.a, c1.a};
            const C aC[6] = {xm.a, xq.a, cur.cym.a, cyp.a, czm.a, cur.czp.a};
            const C s0 = primal_apply<float, FORM_DIFF>(dC, aC, c0.a, vA0 + vz0, H);
            const C dZ[8] = {c0.a, xq.a, cyp.a, e1.a, cur.czp.a, e2.a, cur.v3.a, e3.a};
            const C aZ[6] = {w3.a, zxp.a, cur.u2.a, cur.zyp.a, u1.a, cur.zzp.a};
            const C s1 = primal_apply<float, FORM_DIFF>(dZ, aZ, c1.a, vA1 + vz1, H);
            if (MODE == MODE_REAL) {
                o0.a.x = cur.n0.a.x + s0.y;
                o0.a.y = cur.n0.a.y + (-s0.x);
                o1.a.x = cur.n1.a.x + s1.y;
                o1.a.y = cur.n1.a.y + (-s1.x);
            } else {
                o0.a.x = c0.a.x - s0.x;
                o0.a.y = c0.a.y - s0.y;
                o1.a.x = c1.a.x - s1.x;
                o1.a.y = c1.a.y - s1.y;
            }
        }
        {
            const C dC[8] = {w0.b, u0.b, w1.b, u1.b, w2.b, cur.u2.b, w3.b, c1.b};
            const C aC[6] = {xm.b, xq.b, cur.cym.b, cyp.b, czm.b, cur.czp.b};
            const C s0 = primal_apply<float, FORM_DIFF>(dC, aC, c0.b, vB0 + vz0, H);
            const C dZ[8] = {c0.b, xq.b, cyp.b, e1.b, cur.czp.b, e2.b, cur.v3.b, e3.b};
            const C aZ[6] = {w3.b, zxp.b, cur.u2.b, cur.zyp.b, u1.b, cur.zzp.b};
            const C s1 = primal_apply<float, FORM_DIFF>(dZ, aZ, c1.b, vB1 + vz1, H);
            if (MODE == MODE_REAL) {
                o0.b.x = cur.n0.b.x + s0.y;
                o0.b.y = cur.n0.b.y + (-s0.x);
                o1.b.x = cur.n1.b.x + s1.y;
                o1.b.y = cur.n1.b.y + (-s1.x);
            } else {
                o0.b.x = c0.b.x - s0.x;
                o0.b.y = c0.b.y - s0.y;
                o1.b.x = c1.b.x - s1.x;
                o1.b.y = c1.b.y - s1.y;
            }
        }
        if (!actB) {  // the second cell is the halo column x = nx (or beyond): it stays zero
            o0.b.x = o0.b.y = 0.0f;
            o1.b.x = o1.b.y = 0.0f;
        }
        st_pair<NT || MODE == MODE_IMAG ? kBufNT : 0>(o0, rn, vst, (lay + RS(0, 0, 0)) * ES);
        st_pair<NT || MODE == MODE_IMAG ? kBufNT : 0>(o1, rn, vst, (lay + RS(0, 0, 1)) * ES);
        if (MODE == MODE_IMAG) {
            if (actA)
                part += ((double)o0.a.x * (double)o0.a.x + (double)o0.a.y * (double)o0.a.y) +
                        ((double)o1.a.x * (double)o1.a.x + (double)o1.a.y * (double)o1.a.y);
            if (actB)
                part += ((double)o0.b.x * (double)o0.b.x + (double)o0.b.y * (double)o0.b.y) +
                        ((double)o1.b.x * (double)o1.b.x + (double)o1.b.y * (double)o1.b.y);
        }
        czm = c0;
        c0 = cur.czp;
        u1 = c1;
        c1 = cur.zzp;
        u0 = cur.u2;
        w0 = w2;
        w1 = w3;
        cyp = cur.v3;
        e1 = e3;
        xq = e2;
    };
    if constexpr (CZ > 0) {
#pragma unroll
        for (int kk = 0; kk < CZ; kk++, lay += sz) {
            if (k0 + kk >= k1) break;
            step(k0 + kk, lay, buf[kk & 1], buf[(kk + 1) & 1]);
        }
    } else {
        // a run-time chunk: pairs of layers (fixed registers per buffer), then the rest
        int k = k0;
        for (; k + 1 < k1; k += 2, lay += 2 * sz) {
            step(k, lay, buf[0], buf[1]);
            step(k + 1, lay + sz, buf[1], buf[0]);
        }
        if (k < k1) step(k, lay, buf[0], buf[1]);
    }
    return part;
}

// W > 0 holds the compiler to W waves per SIMD (at most 512 / W registers
// per lane) instead of its own trade of occupancy for registers.
template <int MODE, int TY, int CZ, bool NT, int LY, int W>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1)))
    k_primal_zp(cplx_t<float> *__restrict__ next, const cplx_t<float> *__restrict__ prev, Layout L, PCoef<float> H,
                PTrap trap, double dt, const float *__restrict__ vzt, const double *__restrict__ scale,
                double *__restrict__ partials, Grid3 G) {
    __shared__ double lds[TY];
    int bx, by, bz;
    block_coords(G, bx, by, bz);
    double part[1];
    part[0] = primal_column_p<MODE, TY, CZ, NT, LY>(next, prev, L, H, trap, dt, vzt, scale, bx, by, bz, G.cz);
    if (MODE == MODE_IMAG) {
        block_sum<1, TY>(part, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[bx + G.gx * (by + G.gy * bz)] = part[0];
    }
}

// FUSED (real time): one pass per leapfrog step, out of place,
//   O' = O - i dt H(E) E,   E' = E - i dt H(O') O',
// reading E and O once and writing E' and O' once (64 B per site-step in
// FP64 instead of 96 B for two passes).  A block of TY waves marches a
// z-chunk of CZ layers over a footprint of 64 x TY columns whose interior
// 62 x (TY - 2) columns it owns.  At layer t every footprint lane computes
// O'(t) exactly as primal_column's layer step does (the E rows carried and
// prefetched, x-shifts by DPP), stores it if it owns the column, and puts it
// in a 4-slot LDS ring; after one barrier each interior lane computes
// E'(t - 1) from O' at layers t-2..t of its own and the 8 surrounding columns
// (all read from LDS) and E(t - 1) still in registers.  The ring columns and
// the layers k0 - 1 and k1 of O' are computed again by the neighbouring
// block; O' values outside the grid are zero (the halo).  Same arithmetic and
// neighbour order as the two passes: bit-identical.
template <typename T, int FORM, int TY, int CZ>
__global__ void __launch_bounds__(64 * TY)
    k_primal_fused(cplx_t<T> *__restrict__ eout, cplx_t<T> *__restrict__ oout, const cplx_t<T> *__restrict__ ein,
                   const cplx_t<T> *__restrict__ oin, Layout L, PCoef<T> H, PTrap trap, double dt, Grid3 G) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    __shared__ C ring[4][TY][64][2];
    int bx, by, bz;
    block_coords(G, bx, by, bz);
    const int lane = threadIdx.x, w = threadIdx.y;
    const int i = bx * 62 - 1 + lane, j = by * (TY - 2) - 1 + w;
    const int k0 = bz * CZ, k1 = min(k0 + CZ, L.nz);
    const bool ingrid = i >= 0 && i < L.nx && j >= 0 && j < L.ny;
    const bool owner = ingrid && lane >= 1 && lane <= 62 && w >= 1 && w <= TY - 2;
    const bool left_end = ingrid && lane == 0, right_end = ingrid && lane == 63;
    const long long xp = L.xp, sy = L.sy, sz = L.sz;
    // lanes outside the grid read the zero halo column x = -1 of an existing row
    const int ie = ingrid ? i : -1;
    const int je = ingrid ? j : min(max(j, 0), L.ny - 1);
    const long long lane0 = L.origin + (long long)je * sy + ie;  // corner site of layer 0
    C zero;
    zero.x = (T)0;
    zero.y = (T)0;
    // E at element lane0 + u of layer `layer`; layers beyond the halo are zero
    // (the zero is built here, not captured: a captured local would live in
    // scratch memory)
    auto ldE = [&](long long u, int layer) -> C {
        C v;
        v.x = (T)0;
        v.y = (T)0;
        if (layer >= -1 && layer <= L.nz) v = ein[lane0 + u];
        return v;
    };
    auto ldEe = [&](long long uL, long long uR, int layer) -> C {  // row-end x-neighbours
        C v;
        v.x = (T)0;
        v.y = (T)0;
        if (layer >= -1 && layer <= L.nz) v = ein[lane0 + (left_end ? uL : right_end ? uR : 0)];
        return v;
    };
    auto fetch = [&](int t, ZLayer<T> &f) {  // loads for the O' step of layer t
        const long long u = (long long)t * sz, v = u + xp;
        f.czp = ldE(u + sz, t + 1);
        f.v3 = ldE(u + sz + sy, t + 1);
        f.zzp = ldE(v + sz, t + 1);
        f.cym = ldE(u - sy, t);
        f.u2 = ldE(v - sy, t);
        f.zyp = ldE(v + sy, t);
        f.t0 = left_end ? ldE(v - sy - 1, t) : ldE(u + sz + 1, t + 1);
        f.t1 = left_end ? ldE(v - 1, t) : ldE(u + sz + sy + 1, t + 1);
        f.t2 = ldEe(u - 1, v + 1, t);
        f.n0.x = (T)0;
        f.n0.y = (T)0;
        f.n1 = f.n0;
        if (ingrid && t >= 0 && t < L.nz) {
            f.n0 = oin[lane0 + u];
            f.n1 = oin[lane0 + v];
        }
    };
    int t = k0 - 1;
    long long u = (long long)t * sz;
    C c0 = ldE(u, t), c1 = ldE(u + xp, t), czm = ldE(u - sz, t - 1), u1 = ldE(u + xp - sz, t - 1);
    C u0 = ldE(u + xp - sz - sy, t - 1), cyp = ldE(u + sy, t);
    const C q0 = left_end ? ldE(u + xp - sz - sy - 1, t - 1) : ldE(u + sy + 1, t);  // w0 | e1
    const C q1 = left_end ? ldE(u + xp - sz - 1, t - 1) : ldE(u + 1, t);            // w1 | xq
    ZLayer<T> buf[2];
    fetch(t, buf[0]);
    C w0 = shift_lane_c<true>(u0), w1 = shift_lane_c<true>(u1);
    C e1 = shift_lane_c<false>(cyp), xq = shift_lane_c<false>(c0);
    if (left_end) {
        w0 = q0;
        w1 = q1;
    }
    if (right_end) {
        e1 = q0;
        xq = q1;
    }
    const double xy0 = site_xy(trap, i, j, 0.0), xy1 = site_xy(trap, i, j, 0.5);
    // a fixed trip count with a block-uniform guard (no break), so the loop
    // unrolls despite the barrier and buf[kk & 1] names fixed registers
#pragma unroll
    for (int kk = 0; kk < CZ + 2; kk++, t++, u += sz) {
      if (t <= k1) {
        ZLayer<T> &cur = buf[kk & 1];
        if (t + 1 <= k1) fetch(t + 1, buf[(kk + 1) & 1]);
        __builtin_amdgcn_sched_barrier(0);
        // ---- O'(t) on the footprint ----
        C w2 = shift_lane_c<true>(cur.u2), w3 = shift_lane_c<true>(c1), xm = shift_lane_c<true>(c0);
        C e2 = shift_lane_c<false>(cur.czp), e3 = shift_lane_c<false>(cur.v3), zxp = shift_lane_c<false>(c1);
        if (left_end) {
            w2 = cur.t0;
            w3 = cur.t1;
            xm = cur.t2;
        }
        if (right_end) {
            e2 = cur.t0;
            e3 = cur.t1;
            zxp = cur.t2;
        }
        C o0 = zero, o1 = zero;
        if (ingrid && t >= 0 && t < L.nz) {
            const C dC[8] = {w0, u0, w1, u1, w2, cur.u2, w3, c1};
            const C aC[6] = {xm, xq, cur.cym, cyp, czm, cur.czp};
            const C s0 = primal_apply<T, FORM>(dC, aC, c0, (T)(site_v_xy(trap, xy0, t, 0.0) * dt), H);
            const C dZ[8] = {c0, xq, cyp, e1, cur.czp, e2, cur.v3, e3};
            const C aZ[6] = {w3, zxp, cur.u2, cur.zyp, u1, cur.zzp};
            const C s1 = primal_apply<T, FORM>(dZ, aZ, c1, (T)(site_v_xy(trap, xy1, t, 0.5) * dt), H);
            o0.x = cur.n0.x + s0.y;
            o0.y = cur.n0.y + (-s0.x);
            o1.x = cur.n1.x + s1.y;
            o1.y = cur.n1.y + (-s1.x);
            if (owner && t >= k0 && t < k1) {
                oout[lane0 + u] = o0;
                oout[lane0 + u + xp] = o1;
            }
        }
        ring[t & 3][w][lane][0] = o0;
        ring[t & 3][w][lane][1] = o1;
        __syncthreads();
        // ---- E'(t-1) on the interior, O' from the ring ----
        const int te = t - 1;
        if (owner && te >= k0 && te < k1) {
            const int sm = (te - 1) & 3, s0i = te & 3, sp = (te + 1) & 3;
            C dC[8], aC[6], dZ[8], aZ[6];
            // corner site: centres at dz, dy, dx in {-1, 0}; centre site: corners at {0, +1}
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++)
#pragma unroll
                    for (int dx = 0; dx < 2; dx++) {
                        dC[4 * dz + 2 * dy + dx] = ring[dz ? s0i : sm][w - 1 + dy][lane - 1 + dx][1];
                        dZ[4 * dz + 2 * dy + dx] = ring[dz ? sp : s0i][w + dy][lane + dx][0];
                    }
            aC[0] = ring[s0i][w][lane - 1][0];
            aC[1] = ring[s0i][w][lane + 1][0];
            aC[2] = ring[s0i][w - 1][lane][0];
            aC[3] = ring[s0i][w + 1][lane][0];
            aC[4] = ring[sm][w][lane][0];
            aC[5] = ring[sp][w][lane][0];
            aZ[0] = ring[s0i][w][lane - 1][1];
            aZ[1] = ring[s0i][w][lane + 1][1];
            aZ[2] = ring[s0i][w - 1][lane][1];
            aZ[3] = ring[s0i][w + 1][lane][1];
            aZ[4] = ring[sm][w][lane][1];
            aZ[5] = ring[sp][w][lane][1];
            const C p0 = ring[s0i][w][lane][0], p1 = ring[s0i][w][lane][1];
            const C s0 = primal_apply<T, FORM>(dC, aC, p0, (T)(site_v_xy(trap, xy0, te, 0.0) * dt), H);
            const C s1 = primal_apply<T, FORM>(dZ, aZ, p1, (T)(site_v_xy(trap, xy1, te, 0.5) * dt), H);
            // E(te) is the layer below this step's: czm (corner) and u1 (centre)
            C r0, r1;
            r0.x = czm.x + s0.y;
            r0.y = czm.y + (-s0.x);
            r1.x = u1.x + s1.y;
            r1.y = u1.y + (-s1.x);
            eout[lane0 + u - sz] = r0;
            eout[lane0 + u - sz + xp] = r1;
        }
        // layer t+1 from what layer t read
        czm = c0;
        c0 = cur.czp;
        u1 = c1;
        c1 = cur.zzp;
        u0 = cur.u2;
        w0 = w2;
        w1 = w3;
        cyp = cur.v3;
        e1 = e3;
        xq = e2;
      }
    }
}

// Block partials of {sum|psi|^2, sum 1/2 Re(conj(psi)(-L psi)), sum V|psi|^2,
// sum g|psi|^4} over both sites of every cell, FP64, unscaled weights.
constexpr int kObsTY = 4;
template <typename T>
__global__ void __launch_bounds__(64 * kObsTY) k_primal_observables(const cplx_t<T> *__restrict__ psi, Layout L,
                                                                    PTrap trap, double lD, double lA, double g,
                                                                    double *__restrict__ partials) {
    __shared__ double lds[4 * kObsTY];
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * kObsTY + threadIdx.y;
    const int k = blockIdx.z;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < L.nx && j < L.ny) {
        const long long base = L.origin + (long long)k * L.sz + (long long)j * L.sy + i;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const long long at = base + s * L.xp;
            const cplx_t<T> c = psi[at];
            const double pr = (double)c.x, pi = (double)c.y;
            const int o = s == 0 ? -1 : 0;
            const long long other = base + (1 - s) * L.xp;
            double lr = 0.0, li = 0.0;
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++)
#pragma unroll
                    for (int dx = 0; dx < 2; dx++) {
                        const cplx_t<T> n = psi[other + (o + dx) + (o + dy) * L.sy + (o + dz) * L.sz];
                        lr += lD * ((double)n.x - pr);
                        li += lD * ((double)n.y - pi);
                    }
            const long long ax[6] = {-1, 1, -L.sy, L.sy, -L.sz, L.sz};
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const cplx_t<T> n = psi[at + ax[q]];
                lr += lA * ((double)n.x - pr);
                li += lA * ((double)n.y - pi);
            }
            const double nsq = pr * pr + pi * pi;
            acc[0] += nsq;
            acc[1] += 0.5 * -(pr * lr + pi * li);
            acc[2] += site_v(trap, i, j, k, s ? 0.5 : 0.0) * nsq;
            acc[3] += g * nsq * nsq;
        }
    }
    block_sum<4, kObsTY>(acc, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
#pragma unroll
        for (int q = 0; q < 4; q++) partials[4 * b + q] = acc[q];
    }
}

// Tile shapes: waves per block, block order over the XCDs (block_coords),
// non-temporal read-modify-write, x-neighbour exchange (XS_*), and layers per
// block (0: one layer per block, primal_cell; > 0: a z-march, primal_column,
// which always exchanges by DPP; -1: a buffer z-march whose chunks are sized
// at launch, runtime_chunks).
struct PTile {
    int ty, remap, nt, xs, cz;
    int ly;     // z-march: 1 y-neighbour rows loaded in the layer instead of prefetched, 2 (FP32 two
                // cells per lane) the row-end loads as well
    int fused;  // real time: both half-updates in one z-march pass (k_primal_fused)
    int buf;    // z-march with buffer-resource addressing (k_primal_zb; fields below 2 GiB):
                // the layers of prefetch (1 or 2)
    int pair;   // FP32: two cells per lane (k_primal_zp); FP64 runs k_primal_zb of the same shape
    int waves;  // FP32 two cells per lane: waves per SIMD the registers must allow (0: the compiler's choice)
};
constexpr PTile kPrimalTiles[] = {
    {4, 2, 1, XS_DPP, 0},  {4, 2, 1, XS_LOAD, 0}, {2, 2, 1, XS_DPP, 0},  {8, 2, 1, XS_DPP, 0},
    {2, 2, 1, XS_LOAD, 0}, {8, 2, 1, XS_LOAD, 0}, {4, 0, 1, XS_DPP, 0},  {4, 1, 1, XS_DPP, 0},
    {4, 2, 0, XS_DPP, 0},  {4, 2, 1, XS_SHFL, 0}, {4, 2, 0, XS_LOAD, 0}, {4, 2, 1, XS_DPP, 8},
    {4, 2, 1, XS_DPP, 16}, {4, 2, 1, XS_DPP, 4},  {2, 2, 1, XS_DPP, 8},  {8, 2, 1, XS_DPP, 8},
    {4, 0, 1, XS_DPP, 8},  {4, 2, 0, XS_DPP, 8},  {4, 2, 1, XS_DPP, 16, 1}, {4, 2, 1, XS_DPP, 8, 1},
    {8, 2, 0, XS_DPP, 16, 0, 1}, {8, 2, 0, XS_DPP, 8, 0, 1}, {6, 2, 0, XS_DPP, 16, 0, 1},
    {8, 0, 0, XS_DPP, 16, 0, 1}, {4, 2, 0, XS_DPP, 16, 0, 1},
    {4, 2, 1, XS_DPP, 16, 1, 0, 1}, {4, 2, 1, XS_DPP, 8, 1, 0, 1}, {4, 2, 1, XS_DPP, 16, 0, 0, 1},
    {2, 2, 1, XS_DPP, 16, 1, 0, 1}, {8, 2, 1, XS_DPP, 16, 1, 0, 1}, {4, 2, 0, XS_DPP, 16, 1, 0, 1},
    {4, 2, 1, XS_DPP, 16, 1, 0, 2}, {4, 2, 1, XS_DPP, 16, 0, 0, 2}, {2, 2, 1, XS_DPP, 16, 1, 0, 2},
    {4, 2, 0, XS_DPP, 16, 1, 0, 2}, {8, 2, 1, XS_DPP, 16, 1, 0, 2},
    {4, 2, 1, XS_DPP, 16, 0, 0, 1, 1}, {4, 2, 1, XS_DPP, 16, 1, 0, 1, 1}, {2, 2, 1, XS_DPP, 16, 0, 0, 1, 1},
    {8, 2, 1, XS_DPP, 16, 0, 0, 1, 1}, {4, 2, 0, XS_DPP, 16, 0, 0, 1, 1}, {4, 2, 1, XS_DPP, 8, 0, 0, 1, 1},
    {4, 2, 1, XS_DPP, 16, 1, 0, 1, 1, 4}, {4, 2, 1, XS_DPP, 16, 0, 0, 1, 1, 4}, {4, 2, 0, XS_DPP, 8, 0, 0, 1, 1},
    {4, 2, 0, XS_DPP, 4, 0, 0, 1, 1}, {2, 2, 0, XS_DPP, 8, 0, 0, 1, 1}, {4, 2, 0, XS_DPP, 8, 1, 0, 1, 1, 4},
    {4, 2, 1, XS_DPP, -1, 0, 0, 1, 1}, {4, 2, 1, XS_DPP, -1, 1, 0, 1, 1}, {2, 2, 1, XS_DPP, -1, 0, 0, 1, 1},
    {8, 2, 1, XS_DPP, -1, 0, 0, 1, 1}, {4, 2, 0, XS_DPP, -1, 0, 0, 1, 1}, {4, 2, 1, XS_DPP, -1, 0, 0, 2},
    {4, 2, 0, XS_DPP, 8, 2, 0, 1, 1}, {4, 2, 0, XS_DPP, 8, 2, 0, 1, 1, 4}, {4, 2, 1, XS_DPP, 16, 2, 0, 1, 1, 4},
    {8, 2, 0, XS_DPP, 8, 0, 0, 1, 1}, {4, 1, 0, XS_DPP, 8, 0, 0, 1, 1}, {4, 1, 0, XS_DPP, 16, 0, 0, 1, 1},
    {2, 1, 0, XS_DPP, 8, 0, 0, 1, 1}, {4, 1, 0, XS_DPP, 4, 0, 0, 1, 1}, {8, 1, 0, XS_DPP, 8, 0, 0, 1, 1},
    {2, 1, 0, XS_DPP, 16, 0, 0, 1, 1}, {4, 1, 1, XS_DPP, 8, 0, 0, 1, 1}};
constexpr int kNumPrimalTiles = sizeof(kPrimalTiles) / sizeof(kPrimalTiles[0]);
constexpr int kDefaultPrimalTile = 12;  // {4, 2, 1, XS_DPP, 16}: fields of 2 GiB or more, FP64
constexpr int kMarchTileFP32 = 18;      // {4, 2, 1, XS_DPP, 16, 1}: fields of 2 GiB or more, FP32
// fields the Infinity Cache holds: cached read-modify-write, each XCD a
// contiguous slab of layers (its L2 keeps the slab's halo rows)
constexpr int kCachedTileFP64 = 59;     // {4, 1, 0, XS_DPP, 16, 0, 0, 1, 1}: buffer z-march, one cell per lane
constexpr int kCachedTileFP32 = 59;     // the same shape, two cells per lane
constexpr int kBufMarchTileFP64 = 27;   // {4, 2, 1, XS_DPP, 16, 0, 0, 1}
constexpr int kBufMarchTileFP32 = 36;   // {4, 2, 1, XS_DPP, 16, 0, 0, 1, 1}: two cells per lane

Layout primal_layout(uint32_t nx, uint32_t ny, uint32_t nz, size_t elem_bytes) {
    Layout L{};
    L.nx = (int)nx;
    L.ny = (int)ny;
    L.nz = (int)nz;
    const long long line = (long long)(128 / elem_bytes);  // elements per 128-byte line
    L.line = (int)line;
    // as the dual layout: x = nx and, in the row before, x = -1 share the
    // row's last padding line; one line in front of the first row
    L.xp = (((long long)nx + 1 + line - 1) / line) * line;
    L.sy = 2LL * L.xp;
    L.sz = (long long)(ny + 2) * L.sy;
    L.origin = L.sz + L.sy + line;
    L.elems = (long long)(nz + 2) * L.sz + line;
    return L;
}

Grid3 primal_grid(const PTile &t, const Layout &L) {
    Grid3 G{};
    if (t.pair && L.line == 16) {
        // FP32 (16 elements per 128-byte line) two cells per lane: 128-cell rows
        G.lsh = 6;
        G.gx = (L.nx + 127) / 128;
        G.gy = (L.ny + t.ty - 1) / t.ty;
        G.gz = t.cz > 0 ? (L.nz + t.cz - 1) / t.cz : L.nz;  // run-time chunks: at most one layer each
        G.remap = t.remap;
        if (G.remap == 2 && G.gy % 8 != 0) G.remap = 0;
        return G;
    }
    if (t.fused) {
        // 64-lane rows, interior 62 x (ty - 2) columns per block
        G.lsh = 6;
        G.gx = (L.nx + 61) / 62;
        G.gy = (L.ny + t.ty - 3) / (t.ty - 2);
        G.gz = (L.nz + t.cz - 1) / t.cz;
        G.remap = t.remap;
        if (G.remap == 2 && G.gy % 8 != 0) G.remap = 0;
        return G;
    }
    G.lsh = row_lanes_log2(L.nx);
    const int lw = 1 << G.lsh, rows = t.ty * (64 >> G.lsh);
    G.gx = (L.nx + lw - 1) / lw;
    G.gy = (L.ny + rows - 1) / rows;
    G.gz = t.cz > 0 ? (L.nz + t.cz - 1) / t.cz : L.nz;  // one layer per block, or run-time chunks (at most)
    G.remap = t.remap;
    if (G.remap == 2 && G.gy % 8 != 0) G.remap = 0;  // y-bands need gy divisible by the 8 XCDs
    return G;
}

// Run-time z-chunks (tile cz < 0): as many chunks per column as make one
// round of blocks fill every block slot of the chip (the kernel's own
// occupancy), so that no second, partial round of waves runs, and chunks as
// long as that allows (fewer chunk boundaries, whose halo layers are read
// twice).  Sets G.gz and G.cz.
template <typename K>
hipError_t runtime_chunks(K kern, int ty, const Layout &L, Grid3 &G) {
    static int blocks_per_cu = -1;  // per kernel instantiation: a property of the compiled kernel
    if (blocks_per_cu < 0) {
        int nb = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 64 * ty, 0);
        if (e != hipSuccess) return e;
        blocks_per_cu = nb > 0 ? nb : 1;
    }
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const long long slots = (long long)blocks_per_cu * cus, cols = (long long)G.gx * G.gy;
    const int ntz = (int)std::max(1LL, std::min<long long>(L.nz, slots / cols));
    G.cz = (L.nz + ntz - 1) / ntz;
    G.gz = (L.nz + G.cz - 1) / G.cz;
    return hipSuccess;
}

template <typename T, int FORM, int MODE, int TY, bool NT, int XSM>
void launch_t(cplx_t<T> *next, const cplx_t<T> *prev, const Layout &L, const PCoef<T> &H, const PTrap &trap,
              double dt, const double *scale, double *partials, const Grid3 &G, hipStream_t st) {
    hipLaunchKernelGGL((k_primal<T, FORM, MODE, TY, NT, XSM>), dim3(G.gx * G.gy * G.gz), dim3(64, TY), 0, st, next,
                       prev, L, H, trap, dt, scale, partials, G);
}

// Instantiates one kernel per tile shape (imaginary time always stores
// streaming, so its shapes differ only in waves, order and exchange).
// *nblocks (if given) receives the number of blocks launched, whose partial
// sums an imaginary-time pass leaves in partials[0 .. nblocks).
template <typename T, int FORM, int MODE, int IDX = 0>
hipError_t launch_primal(int tile, cplx_t<T> *next, const cplx_t<T> *prev, const Layout &L, const PCoef<T> &H,
                         const PTrap &trap, double dt, const double *scale, double *partials, hipStream_t st,
                         bool reverse, const float *vzt, long long *nblocks = nullptr) {
    if constexpr (IDX < kNumPrimalTiles) {
        if (tile != IDX)
            return launch_primal<T, FORM, MODE, IDX + 1>(tile, next, prev, L, H, trap, dt, scale, partials, st,
                                                         reverse, vzt, nblocks);
        constexpr PTile t = kPrimalTiles[IDX];
        if constexpr (t.fused) {
            return hipErrorInvalidValue;  // launched by launch_primal_fused
        }
        Grid3 G = primal_grid(t, L);
        G.rev = reverse ? 1 : 0;
        constexpr bool nt = MODE == MODE_REAL && t.nt;
        constexpr int czt = t.cz < 0 ? 0 : t.cz;  // 0: chunks sized at launch (buffer z-marches only)
        static_assert(t.cz >= 0 || t.buf, "run-time z-chunks need a buffer-addressed z-march");
        if constexpr (t.pair && sizeof(T) == 4) {
            const auto kern = k_primal_zp<MODE, t.ty, czt, nt, t.ly, t.waves>;
            if constexpr (t.cz < 0) {
                const hipError_t e = runtime_chunks(kern, t.ty, L, G);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(kern, dim3(G.gx * G.gy * G.gz), dim3(64, t.ty), 0, st, next, prev, L, H, trap, dt, vzt,
                               scale, partials, G);
        } else if constexpr (t.buf) {
            const auto kern = k_primal_zb<T, FORM, MODE, t.ty, czt, nt, t.ly != 0, t.buf>;
            if constexpr (t.cz < 0) {
                const hipError_t e = runtime_chunks(kern, t.ty, L, G);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(kern, dim3(G.gx * G.gy * G.gz), dim3(64, t.ty), 0, st, next, prev, L, H, trap, dt, vzt,
                               scale, partials, G);
        } else if constexpr (t.cz > 0)
            hipLaunchKernelGGL((k_primal_zm<T, FORM, MODE, t.ty, t.cz, nt, t.ly != 0>), dim3(G.gx * G.gy * G.gz), dim3(64, t.ty),
                               0, st, next, prev, L, H, trap, dt, scale, partials, G);
        else
            launch_t<T, FORM, MODE, t.ty, nt, t.xs>(next, prev, L, H, trap, dt, scale, partials, G, st);
        if (nblocks) *nblocks = (long long)G.gx * G.gy * G.gz;
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
}

// One fused leapfrog step (real time) with a fused tile shape.
template <typename T, int FORM, int IDX = 0>
hipError_t launch_primal_fused(int tile, cplx_t<T> *eout, cplx_t<T> *oout, const cplx_t<T> *ein,
                               const cplx_t<T> *oin, const Layout &L, const PCoef<T> &H, const PTrap &trap, double dt,
                               hipStream_t st) {
    if constexpr (IDX < kNumPrimalTiles) {
        if (tile != IDX)
            return launch_primal_fused<T, FORM, IDX + 1>(tile, eout, oout, ein, oin, L, H, trap, dt, st);
        constexpr PTile t = kPrimalTiles[IDX];
        if constexpr (!t.fused) {
            return hipErrorInvalidValue;
        } else {
            const Grid3 G = primal_grid(t, L);
            hipLaunchKernelGGL((k_primal_fused<T, FORM, t.ty, t.cz>), dim3(G.gx * G.gy * G.gz), dim3(64, t.ty), 0, st,
                               eout, oout, ein, oin, L, H, trap, dt, G);
            return hipGetLastError();
        }
    } else {
        return hipErrorInvalidValue;
    }
}

long long primal_blocks(int tile, const Layout &L) {
    const Grid3 G = primal_grid(kPrimalTiles[tile], L);
    return (long long)G.gx * G.gy * G.gz;
}

long long primal_obs_blocks(const Layout &L) {
    return (long long)((L.nx + 63) / 64) * ((L.ny + kObsTY - 1) / kObsTY) * L.nz;
}

double primal_vmax(const uint32_t dims[3], const double origin[3], double a, double ratio) {
    // sites span p0 + a [0, n - 1/2] per axis; the extreme |coordinate| bounds V
    double m[3];
    for (int d = 0; d < 3; d++) m[d] = std::max(std::fabs(origin[d]), std::fabs(origin[d] + a * (dims[d] - 0.5)));
    return 0.5 * (m[0] * m[0] + m[1] * m[1] + ratio * ratio * m[2] * m[2]);
}

// -L/2 has spectrum in [0, 7.5/a^2] (Gershgorin: diagonal 15/(8a^2), off-diagonal
// row sum 15/(8a^2)), so both integrators need dt (7.5/a^2 + max V) < 2.
bool primal_stable(double dt, double a, double vmax) { return dt * (7.5 / (a * a) + vmax) < 2.0; }

}  // namespace

struct gpe_primal {
    gpe_primal_config cfg{};
    Layout L{};
    PTrap trap{};
    double coef[4] = {0, 0, 0, 0};
    double w = 0.0;  // site measure a^3/2
    int tile = 0;
    int imag_tile = 0;  // the size's default shape, for imaginary time under a fused tile
    // field[ie] even level, field[io] odd level / imaginary-time scratch; a
    // fused tile writes a step into the other two and swaps
    void *field[4] = {nullptr, nullptr, nullptr, nullptr};
    int ie = 0, io = 1;
    double *scal = nullptr;      // [0] last N, [1] read scale, [4..7] observables
    double *partials = nullptr;  // per-block partial sums
    long long npartials = 0;
    // FP32: (0.5 rr z^2 dt) per layer and sublattice, rounded to FP32, for the
    // buffer-addressed z-march (recomputed when dt changes)
    float *vzt = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t steps_done = 0;
    bool odd_valid = false;

    ~gpe_primal() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto *p : field)
            if (p) (void)hipFree(p);
        if (scal) (void)hipFree(scal);
        if (partials) (void)hipFree(partials);
        if (vzt) (void)hipFree(vzt);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    size_t elem_bytes() const { return cfg.precision == GPE_FP64 ? 16 : 8; }
    size_t sites() const { return 2ull * cfg.nx * cfg.ny * cfg.nz; }
    void set_coefs() {
        gpe_primal_coefficients(cfg.a, cfg.dt, cfg.g, coef);
        w = 0.5 * cfg.a * cfg.a * cfg.a;
    }
    gpe_status init();
    gpe_status upload_vz();
    template <typename T> gpe_status upload(int f, const double *h);
    template <typename T> gpe_status download(int f, double *h);
    template <typename T, int FORM> gpe_status enqueue(uint64_t nsteps, uint64_t *launches);
    template <typename T> gpe_status observables_t(int f, double out[4]);
};

gpe_status gpe_primal::init() {
    L = primal_layout(cfg.nx, cfg.ny, cfg.nz, elem_bytes());
    for (int d = 0; d < 3; d++) trap.p0[d] = cfg.origin[d];
    trap.a = cfg.a;
    trap.rr = cfg.trap_ratio * cfg.trap_ratio;
    // default by size, the fastest shapes measured on MI355X: when both fields
    // fit the 256 MB Infinity Cache (128^3 cells) a cached read-modify-write
    // in the buffer-addressed z-march, each XCD a contiguous slab, 16 layers
    // per chunk (FP32: two cells per lane); larger grids (256^3) the
    // buffer-addressed z-march of 16 layers with one layer of prefetch (FP32:
    // two cells per lane); fields of 2 GiB or more (beyond 32-bit offsets) the
    // z-march on 64-bit addresses
    const double fbytes_total = (double)L.elems * (double)elem_bytes();
    const bool cached = 2.0 * fbytes_total < 192e6;
    const bool small = fbytes_total < 2147483648.0;
    const bool fp64 = cfg.precision == GPE_FP64;
    if (cached) imag_tile = fp64 ? kCachedTileFP64 : kCachedTileFP32;
    else if (small) imag_tile = fp64 ? kBufMarchTileFP64 : kBufMarchTileFP32;
    else imag_tile = fp64 ? kDefaultPrimalTile : kMarchTileFP32;
    tile = cfg.tile != 0 ? (int)cfg.tile - 1 : imag_tile;
    if (kPrimalTiles[tile].buf && (double)L.elems * (double)elem_bytes() >= 2147483648.0) {
        set_error("gpe_primal_create: tile %d addresses fields by 32-bit offsets; this grid's fields exceed 2 GiB",
                  tile + 1);
        return GPE_EINVAL;
    }
    set_coefs();
    GPE_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    GPE_HIP_TRY(hipEventCreate(&ev0));
    GPE_HIP_TRY(hipEventCreate(&ev1));
    const size_t fbytes = (size_t)L.elems * elem_bytes();
    const int nfields = kPrimalTiles[tile].fused ? 4 : 2;
    for (int f = 0; f < nfields; f++) {
        if (hipMalloc(&field[f], fbytes) != hipSuccess) {
            (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
            field[f] = nullptr;
            set_error("gpe_primal_create: cannot allocate %zu bytes of device memory", fbytes);
            return GPE_ENOMEM;
        }
        // the zero halo is never written afterwards
        GPE_HIP_TRY(hipMemsetAsync(field[f], 0, fbytes, stream));
    }
    npartials = primal_obs_blocks(L);
    for (int t = 0; t < kNumPrimalTiles; t++) npartials = std::max(npartials, primal_blocks(t, L));
    GPE_HIP_TRY(hipMalloc(&partials, (size_t)npartials * 4 * sizeof(double)));
    GPE_HIP_TRY(hipMalloc(&scal, 8 * sizeof(double)));
    const double init_scal[8] = {1.0, 1.0, 0, 0, 0, 0, 0, 0};
    GPE_HIP_TRY(hipMemcpyAsync(scal, init_scal, sizeof(init_scal), hipMemcpyHostToDevice, stream));
    GPE_HIP_TRY(hipStreamSynchronize(stream));
    if (!fp64) {
        GPE_HIP_TRY(hipMalloc(&vzt, (size_t)L.nz * 2 * sizeof(float)));
        return upload_vz();
    }
    return GPE_OK;
}

// The z part of V*dt per layer k and sublattice s (z = p0 + a (k + s/2)), in
// site_v's operation order, rounded to FP32.
gpe_status gpe_primal::upload_vz() {
    if (!vzt) return GPE_OK;
    std::vector<float> h((size_t)L.nz * 2);
    for (int k = 0; k < L.nz; k++)
        for (int sl = 0; sl < 2; sl++) {
            const double z = trap.p0[2] + trap.a * ((double)k + (sl ? 0.5 : 0.0));
            h[2 * k + sl] = (float)(0.5 * (trap.rr * z * z) * cfg.dt);
        }
    GPE_HIP_TRY(hipMemcpy(vzt, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    return GPE_OK;
}

// Host [k][j][i][s] complex double <-> device class rows.
template <typename T>
gpe_status gpe_primal::upload(int f, const double *h) {
    std::vector<T> d((size_t)L.elems * 2, T(0));
    size_t n = 0;
    for (int k = 0; k < L.nz; k++)
        for (int j = 0; j < L.ny; j++)
            for (int i = 0; i < L.nx; i++)
                for (int s = 0; s < 2; s++, n++) {
                    const long long e = L.origin + (long long)k * L.sz + (long long)j * L.sy + s * L.xp + i;
                    d[2 * e] = (T)h[2 * n];
                    d[2 * e + 1] = (T)h[2 * n + 1];
                }
    GPE_HIP_TRY(hipMemcpyAsync(field[f], d.data(), d.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    GPE_HIP_TRY(hipStreamSynchronize(stream));
    return GPE_OK;
}

template <typename T>
gpe_status gpe_primal::download(int f, double *h) {
    std::vector<T> d((size_t)L.elems * 2);
    GPE_HIP_TRY(hipMemcpyAsync(d.data(), field[f], d.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
    GPE_HIP_TRY(hipStreamSynchronize(stream));
    size_t n = 0;
    for (int k = 0; k < L.nz; k++)
        for (int j = 0; j < L.ny; j++)
            for (int i = 0; i < L.nx; i++)
                for (int s = 0; s < 2; s++, n++) {
                    const long long e = L.origin + (long long)k * L.sz + (long long)j * L.sy + s * L.xp + i;
                    h[2 * n] = (double)d[2 * e];
                    h[2 * n + 1] = (double)d[2 * e + 1];
                }
    return GPE_OK;
}

template <typename T, int FORM>
gpe_status gpe_primal::enqueue(uint64_t nsteps, uint64_t *launches) {
    using C = cplx_t<T>;
    PCoef<T> H;
    H.cD = (T)coef[0];
    H.cA = (T)coef[1];
    H.c0 = (T)coef[2];
    H.g = (T)coef[3];
    const double dt = cfg.dt;
    uint64_t nl = 0;
    hipError_t e = hipSuccess;
    if (cfg.mode == GPE_REAL_TIME && kPrimalTiles[tile].fused) {
        for (uint64_t s = 0; s < nsteps && e == hipSuccess; s++) {
            int spare[2], n = 0;
            for (int f = 0; f < 4; f++)
                if (f != ie && f != io) spare[n++] = f;
            e = launch_primal_fused<T, FORM>(tile, (C *)field[spare[0]], (C *)field[spare[1]], (const C *)field[ie],
                                             (const C *)field[io], L, H, trap, dt, stream);
            if (e != hipSuccess) break;  // nothing was written: the state stays where it is
            ie = spare[0];
            io = spare[1];
            nl += 1;
        }
    } else if (cfg.mode == GPE_REAL_TIME) {
        C *ev = (C *)field[ie], *od = (C *)field[io];
        for (uint64_t s = 0; s < nsteps && e == hipSuccess; s++) {
            // the even update walks the tiles backwards, starting on the layers
            // the odd update touched last
            e = launch_primal<T, FORM, MODE_REAL>(tile, od, ev, L, H, trap, dt, nullptr, nullptr, stream, false, vzt);
            if (e == hipSuccess)
                e = launch_primal<T, FORM, MODE_REAL>(tile, ev, od, L, H, trap, dt, nullptr, nullptr, stream, true, vzt);
            nl += 2;
        }
    } else {
        // Euler step into the other buffer; the renormalisation is folded into
        // the next read through scal[1] = 1/sqrt(N).  (A fused tile is a
        // leapfrog step; imaginary time runs the size's default shape.)
        const int itile = kPrimalTiles[tile].fused ? imag_tile : tile;
        for (uint64_t s = 0; s < nsteps && e == hipSuccess; s++) {
            C *in = (C *)field[ie], *out = (C *)field[io];
            const bool rev = (steps_done + s) & 1;
            long long nb = 0;  // blocks launched, each leaving one partial sum
            e = launch_primal<T, FORM, MODE_IMAG>(itile, out, in, L, H, trap, dt, scal + 1, partials, stream, rev, vzt,
                                                  &nb);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(256), 0, stream, partials, nb, w, 1, scal);
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
        set_error("gpe_primal_step: launch failed: %s", hipGetErrorString(e));
        return GPE_EDEVICE;
    }
    steps_done += nsteps;
    if (launches) *launches = nl;
    return GPE_OK;
}

template <typename T>
gpe_status gpe_primal::observables_t(int f, double out[4]) {
    const dim3 grid((L.nx + 63) / 64, (L.ny + kObsTY - 1) / kObsTY, L.nz), block(64, kObsTY);
    const double a2 = cfg.a * cfg.a;
    hipLaunchKernelGGL((k_primal_observables<T>), grid, block, 0, stream, (const cplx_t<T> *)field[f], L, trap,
                       0.75 / a2, 0.25 / a2, cfg.g, partials);
    GPE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_finalize<4>), dim3(1), dim3(256), 0, stream, partials, primal_obs_blocks(L), w, 0, scal + 4);
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipMemcpyAsync(out, scal + 4, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
    GPE_HIP_TRY(hipStreamSynchronize(stream));
    return GPE_OK;
}

extern "C" {

void gpe_primal_coefficients(double a, double dt, double g, double out[4]) {
    const double a2 = a * a;
    out[0] = (-0.375 / a2) * dt;
    out[1] = (-0.125 / a2) * dt;
    out[2] = (3.75 / a2) * dt;
    out[3] = g * dt;
}

void gpe_primal_centered_origin(const uint32_t dims[3], double a, double origin[3]) {
    for (int d = 0; d < 3; d++) origin[d] = -0.5 * a * ((double)dims[d] - 0.5);
}

void gpe_primal_positions(const uint32_t dims[3], const double origin[3], double a, double *xyz) {
    size_t n = 0;
    for (uint32_t k = 0; k < dims[2]; k++)
        for (uint32_t j = 0; j < dims[1]; j++)
            for (uint32_t i = 0; i < dims[0]; i++)
                for (int s = 0; s < 2; s++, n++) {
                    const double off = s ? 0.5 : 0.0;
                    xyz[3 * n] = origin[0] + a * ((double)i + off);
                    xyz[3 * n + 1] = origin[1] + a * ((double)j + off);
                    xyz[3 * n + 2] = origin[2] + a * ((double)k + off);
                }
}

void gpe_primal_trap_potential(const uint32_t dims[3], const double origin[3], double a, double ratio, double *V) {
    PTrap t{};
    for (int d = 0; d < 3; d++) t.p0[d] = origin[d];
    t.a = a;
    t.rr = ratio * ratio;
    size_t n = 0;
    for (uint32_t k = 0; k < dims[2]; k++)
        for (uint32_t j = 0; j < dims[1]; j++)
            for (uint32_t i = 0; i < dims[0]; i++)
                for (int s = 0; s < 2; s++, n++) V[n] = site_v(t, (int)i, (int)j, (int)k, s ? 0.5 : 0.0);
}

int gpe_primal_tiles(void) { return kNumPrimalTiles; }

gpe_status gpe_primal_create(const gpe_primal_config *cfg, gpe_primal **out) {
    if (!cfg || !out) {
        set_error("gpe_primal_create: null argument");
        return GPE_EINVAL;
    }
    *out = nullptr;
    if (cfg->nx == 0 || cfg->ny == 0 || cfg->nz == 0 || cfg->nx > 65536 || cfg->ny > 65536 || cfg->nz > 65536) {
        set_error("gpe_primal_create: grid %u x %u x %u out of range", cfg->nx, cfg->ny, cfg->nz);
        return GPE_EINVAL;
    }
    if (cfg->precision > GPE_FP32_BULK || cfg->mode > GPE_IMAG_TIME || cfg->tile > (uint32_t)kNumPrimalTiles) {
        set_error("gpe_primal_create: bad precision, mode or tile");
        return GPE_EINVAL;
    }
    if (!(cfg->a > 0) || !(cfg->dt > 0) || !std::isfinite(cfg->a) || !std::isfinite(cfg->dt) ||
        !std::isfinite(cfg->g) || !std::isfinite(cfg->trap_ratio)) {
        set_error("gpe_primal_create: a and dt must be positive, g and trap_ratio finite");
        return GPE_EINVAL;
    }
    gpe_primal_config c = *cfg;
    const uint32_t dims[3] = {c.nx, c.ny, c.nz};
    if (!std::isfinite(c.origin[0]) || !std::isfinite(c.origin[1]) || !std::isfinite(c.origin[2]))
        gpe_primal_centered_origin(dims, c.a, c.origin);
    const double vmax = primal_vmax(dims, c.origin, c.a, c.trap_ratio);
    if (!primal_stable(c.dt, c.a, vmax)) {
        set_error("gpe_primal_create: dt = %g violates dt*(7.5/a^2 + max V) < 2 (max V = %g)", c.dt, vmax);
        return GPE_EUNSTABLE;
    }
    gpe_primal *p = new (std::nothrow) gpe_primal;
    if (!p) {
        set_error("gpe_primal_create: out of host memory");
        return GPE_ENOMEM;
    }
    p->cfg = c;
    const gpe_status st = p->init();
    if (st != GPE_OK) {
        delete p;
        return st;
    }
    *out = p;
    return GPE_OK;
}

void gpe_primal_destroy(gpe_primal *p) { delete p; }

gpe_status gpe_primal_set_state(gpe_primal *p, const double *even, const double *odd) {
    if (!p || !even) {
        set_error("gpe_primal_set_state: null argument");
        return GPE_EINVAL;
    }
    const size_t n = 2 * p->sites();
    for (size_t e = 0; e < n; e++)
        if (!std::isfinite(even[e]) || (odd && !std::isfinite(odd[e]))) {
            set_error("gpe_primal_set_state: non-finite value at %zu", e);
            return GPE_EUNSTABLE;
        }
    GPE_HIP_TRY(hipStreamSynchronize(p->stream));
    const bool fp64 = p->cfg.precision == GPE_FP64;
    gpe_status st = fp64 ? p->upload<double>(p->ie, even) : p->upload<float>(p->ie, even);
    if (st != GPE_OK) return st;
    p->odd_valid = false;
    if (odd) {
        st = fp64 ? p->upload<double>(p->io, odd) : p->upload<float>(p->io, odd);
        if (st != GPE_OK) return st;
        p->odd_valid = true;
    }
    return GPE_OK;
}

gpe_status gpe_primal_get_state(gpe_primal *p, double *even, double *odd) {
    if (!p || !even) {
        set_error("gpe_primal_get_state: null argument");
        return GPE_EINVAL;
    }
    const bool fp64 = p->cfg.precision == GPE_FP64;
    gpe_status st = fp64 ? p->download<double>(p->ie, even) : p->download<float>(p->ie, even);
    if (st == GPE_OK && odd) st = fp64 ? p->download<double>(p->io, odd) : p->download<float>(p->io, odd);
    return st;
}

gpe_status gpe_primal_set_mode(gpe_primal *p, uint32_t mode, double dt) {
    if (!p || mode > GPE_IMAG_TIME || !(dt > 0) || !std::isfinite(dt)) {
        set_error("gpe_primal_set_mode: invalid argument");
        return GPE_EINVAL;
    }
    const uint32_t dims[3] = {p->cfg.nx, p->cfg.ny, p->cfg.nz};
    if (!primal_stable(dt, p->cfg.a, primal_vmax(dims, p->cfg.origin, p->cfg.a, p->cfg.trap_ratio))) {
        set_error("gpe_primal_set_mode: dt = %g violates dt*(7.5/a^2 + max V) < 2", dt);
        return GPE_EUNSTABLE;
    }
    GPE_HIP_TRY(hipStreamSynchronize(p->stream));
    if (mode != p->cfg.mode) p->odd_valid = false;
    p->cfg.mode = mode;
    p->cfg.dt = dt;
    p->set_coefs();
    return p->upload_vz();
}

gpe_status gpe_primal_init_odd(gpe_primal *p, double mu) {
    if (!p || !std::isfinite(mu)) {
        set_error("gpe_primal_init_odd: invalid argument");
        return GPE_EINVAL;
    }
    // getPhase(-dt/2) = (cos(-mu dt/2), -sin(-mu dt/2)) (VortexState.cpp:198-202)
    const double ph = (-0.5 * p->cfg.dt) * mu;
    const double cr = std::cos(ph), ci = -std::sin(ph);
    if (p->cfg.precision == GPE_FP64)
        hipLaunchKernelGGL((k_phase_copy<double>), dim3(1024), dim3(256), 0, p->stream, (double2 *)p->field[p->io],
                           (const double2 *)p->field[p->ie], p->L.elems, cr, ci);
    else
        hipLaunchKernelGGL((k_phase_copy<float>), dim3(1024), dim3(256), 0, p->stream, (float2 *)p->field[p->io],
                           (const float2 *)p->field[p->ie], p->L.elems, cr, ci);
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipStreamSynchronize(p->stream));
    p->odd_valid = true;
    return GPE_OK;
}

static gpe_status primal_step(gpe_primal *p, uint64_t nsteps, uint64_t *launches) {
    if (p->cfg.mode == GPE_REAL_TIME && !p->odd_valid) {
        set_error("gpe_primal_step: real-time stepping needs an odd field (set_state or init_odd)");
        return GPE_EINVAL;
    }
    if (p->cfg.precision == GPE_FP64) return p->enqueue<double, FORM_REF>(nsteps, launches);
    return p->enqueue<float, FORM_DIFF>(nsteps, launches);
}

gpe_status gpe_primal_step(gpe_primal *p, uint64_t nsteps) {
    if (!p) {
        set_error("gpe_primal_step: null solver");
        return GPE_EINVAL;
    }
    return primal_step(p, nsteps, nullptr);
}

gpe_status gpe_primal_time_steps(gpe_primal *p, uint64_t nsteps, double *elapsed_ms, uint64_t *launches) {
    if (!p || !elapsed_ms) {
        set_error("gpe_primal_time_steps: null argument");
        return GPE_EINVAL;
    }
    GPE_HIP_TRY(hipEventRecord(p->ev0, p->stream));
    const gpe_status st = primal_step(p, nsteps, launches);
    if (st != GPE_OK) return st;
    GPE_HIP_TRY(hipEventRecord(p->ev1, p->stream));
    GPE_HIP_TRY(hipEventSynchronize(p->ev1));
    float ms = 0.0f;
    GPE_HIP_TRY(hipEventElapsedTime(&ms, p->ev0, p->ev1));
    *elapsed_ms = ms;
    return GPE_OK;
}

gpe_status gpe_primal_observables(gpe_primal *p, int field, double out[4]) {
    if (!p || !out || (field != 0 && field != 1)) {
        set_error("gpe_primal_observables: invalid argument");
        return GPE_EINVAL;
    }
    const int f = field == 0 ? p->ie : p->io;
    if (p->cfg.precision == GPE_FP64) return p->observables_t<double>(f, out);
    return p->observables_t<float>(f, out);
}

gpe_status gpe_primal_synchronize(gpe_primal *p) {
    if (!p) {
        set_error("gpe_primal_synchronize: null solver");
        return GPE_EINVAL;
    }
    GPE_HIP_TRY(hipStreamSynchronize(p->stream));
    return GPE_OK;
}

}  // extern "C"
