// gpe_fused_row.hpp -- a whole leapfrog step in one z-march pass, one
// independent wavefront per cell row:
//   O' = O - i dt H(E) E        (odd half-update, GPUGrossPitaevskii_BCC.cu:478)
//   E' = E - i dt H(O') O'      (even half-update, GPUGrossPitaevskii_BCC.cu:480)
// out of place, reading E and O once and writing E' and O' once (64 B per
// node-step in FP64 instead of the 96 B of two passes).
//
// Layout of the work.  A wave owns one cell row j and a strip of w <= 62
// x-columns xs .. xs + w - 1 (xs = w bx; w is a launch parameter): lanes
// 1..w are its interior, lanes 0 and w + 1 a one-column ring
// whose O' the interior reads as x-neighbours (DPP wave shifts), lanes
// beyond w + 1 idle.  It marches the layers of a z-chunk (the host sizes the
// chunks so that one round of waves fills the chip): at step t it computes
// O'(t) of its row and E'(t - 1) of its interior.  E'(t - 1) needs O' of the
// rows j - 1 and j + 1 as well, but only some classes of them: the classes
// the row below exports upwards are {5, 6, 7, 8} and the row above exports
// {1, 9} (the y-faces of the stencil, GPUGrossPitaevskii_BCC.cu:93-140).  The
// wave computes those 6 class values of the two neighbouring rows itself,
// from E of those rows and of its own row; on this mesh each of them depends
// only on the row it belongs to and the row between it and the interior, so
// no value beyond rows j - 1 .. j + 1 is read.  Waves therefore never
// exchange data: no LDS, no barriers, and a block is just TY waves of
// neighbouring rows, scheduled together so that the rows they read of each
// other are cache hits.  The same holds for the ring lanes in x (the classes
// x-1 exports, {1, 2, 3, 4}, and x+1 exports, {0, 5}, depend on nothing
// beyond the strip) and for the layers k0 - 1 and k1 that the neighbouring
// z-chunks own.  Values on the zero halo (outside the box) are forced to zero
// rather than computed.
//
// Arithmetic per class is update_class of gpe_fused.hpp: the reference's
// operation sequence, bit-identical to the two-pass kernels in FP64; the
// neighbour values recomputed for the adjacent rows come out of the same
// operations on the same inputs, so they carry the same bits as those rows'
// own results.
#pragma once
#include <type_traits>

#include "gpe_fused.hpp"

namespace gpe {

struct RowGrid {
    int ntx, nty, ntz;  // x-strips, blocks of TY rows, z-chunks
    int w;              // interior columns per strip (at most 62)
    int cz;             // layers per chunk
    int remap;          // XCD-contiguous block order
};

// V*dt of class l of the cell a lane handles, for update_class: the
// in-kernel trap, or the streamed array through the kernel's buffer
// resource (the potential shares the field's node order at half the element
// size, so its byte offsets are the field's halved).
template <typename T, int POT>
struct RowPot {
    __amdgpu_buffer_rsrc_t r;
    uint32_t vlayer;          // per-lane byte offset of the layer (field units)
    const uint32_t *srow;     // wave-uniform class offsets of the row (field units)
    const Trap &trap;
    double dt;
    int i, j, k;
    __device__ __forceinline__ T operator[](int l) const {
        if (POT == POT_TRAP) return (T)(trap_v4(trap, i, j, k, l) * dt);
        const uint32_t s = __builtin_amdgcn_readfirstlane(srow[l] >> 1);
        if constexpr (sizeof(T) == 8)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vlayer >> 1, s, 0));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vlayer >> 1, s, 0));
    }
};

// v where the lane owns a node of the box, 0 elsewhere, selected per
// component (a select of the whole struct can end up as a private-memory
// round trip).
template <typename C>
__device__ __forceinline__ C keep_if(bool keep, C v) {
    v.x = keep ? v.x : (decltype(v.x))0;
    v.y = keep ? v.y : (decltype(v.y))0;
    return v;
}

// FP64 needs about all 256 registers a wave can have at 2 waves per SIMD:
// the compiler is held to that occupancy rather than trading it for more
// loads in flight.
// CZ > 0: layers per chunk fixed at compile time (a constant trip count the
// compiler schedules better); CZ = 0: taken from the grid.
template <typename T, int FORM, int POT, int TY, int CZ>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 2 : 3)))
    k_fused_row(cplx_t<T> *__restrict__ eout, cplx_t<T> *__restrict__ oout, const cplx_t<T> *__restrict__ ein,
                const cplx_t<T> *__restrict__ oin, const T *__restrict__ pot, Layout L, Hamiltonian<T, FORM> H,
                Trap trap, double dt, RowGrid G) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    const int w = remap_block(blockIdx.x, G.ntx * G.nty * G.ntz, G.remap);
    const int bx = w % G.ntx, by = (w / G.ntx) % G.nty, bz = w / (G.ntx * G.nty);
    const int j = by * TY + (int)threadIdx.y;
    if (j >= L.ny) return;  // the whole wave: rows beyond the box
    const int cz = CZ > 0 ? CZ : G.cz;
    const int k0 = bz * cz, k1 = min(k0 + cz, L.nz);
    const int lane = threadIdx.x;
    const int i = bx * G.w - 1 + lane;
    // lanes past the right ring column own nothing and act as zero halo
    const bool dom = i >= 0 && i < L.nx && lane <= G.w + 1;
    const bool inner = dom && lane >= 1 && lane <= G.w;
    const bool ym_ok = j >= 1, yp_ok = j + 1 < L.ny;  // wave-uniform: rows j -+ 1 inside the box
    // lanes outside the box read the zero halo column x = -1 of their row
    const long long col = L.origin + (long long)j * L.sy + (dom ? i : -1);
    C zero;
    zero.x = (T)0;
    zero.y = (T)0;
    // Accesses go through buffer resources.  The per-lane byte offset carries
    // the column and the layer (one add per layer); the wave-uniform one the
    // class and the row, the same for every layer, so the loop holds them in
    // scalar registers.  The host runs this kernel only on fields below 2 GiB.
    constexpr uint32_t ES = sizeof(C);
    const int nbytes = (int)((long long)L.elems * ES);
    const __amdgpu_buffer_rsrc_t rE = __builtin_amdgcn_make_buffer_rsrc((void *)ein, (short)0, nbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rO = __builtin_amdgcn_make_buffer_rsrc((void *)oin, (short)0, nbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t wE = __builtin_amdgcn_make_buffer_rsrc((void *)eout, (short)0, nbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t wO = __builtin_amdgcn_make_buffer_rsrc((void *)oout, (short)0, nbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rP =
        __builtin_amdgcn_make_buffer_rsrc((void *)pot, (short)0, POT == POT_ARRAY ? nbytes / 2 : 0, 0x00020000);
    const uint32_t szb = (uint32_t)L.sz * ES;
    // layer k of this lane's column, row j - 1, class 0 (k in -1..nz: the
    // layers -1 and nz are the zero halo)
    const uint32_t vcol = (uint32_t)(col - L.sz - L.sy) * ES;
    auto vlay = [&](int k) -> uint32_t { return vcol + (uint32_t)(k + 1) * szb; };
    // class l of row j + dy, relative to vlay
    uint32_t srow[3][12];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int l = 0; l < 12; l++)
            srow[d][l] = __builtin_amdgcn_readfirstlane((uint32_t)(l * L.xp + d * L.sy) * ES);
    auto ld = [&](const __amdgpu_buffer_rsrc_t &r, uint32_t v, int l, int dy) -> C {
        if constexpr (sizeof(T) == 8)
            return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(r, v, srow[dy + 1][l], 0));
        else
            return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b64(r, v, srow[dy + 1][l], 0));
    };
    auto st = [&](const __amdgpu_buffer_rsrc_t &r, uint32_t v, int l, C x) {
        if constexpr (sizeof(T) == 8)
            buffer_store_128(x, r, v, srow[1][l]);
        else
            __builtin_amdgcn_raw_buffer_store_b64(
                __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), x), r, v,
                srow[1][l], 0);
    };

    // ---- march state ---------------------------------------------------
    // E(t-1): own row (all classes in FP32, where E'(t-1) takes its base from
    // here; FP64 keeps only classes 0, 9, 10, 11 and reloads the base from the
    // caches, for registers); class 9 of row j+1, class 10 of row j-1
    constexpr bool kCarryE = sizeof(T) == 4;
    C Edn[12];
#pragma unroll
    for (int l = 0; l < 12; l++) Edn[l] = zero;
    C Eyp_dn9 = zero, Eym_dn10 = zero;
    // O'(t-1): own row, the exported classes of rows j-1 {5, 7, 8} and j+1 {1, 9}
    C Op[12];
#pragma unroll
    for (int l = 0; l < 12; l++) Op[l] = zero;
    C Opym5 = zero, Opym7 = zero, Opym8 = zero, Opyp1 = zero, Opyp9 = zero;
    // O'(t-2): own classes 0, 10, 11; class 9 of row j+1
    C Opp0 = zero, Opp10 = zero, Opp11 = zero, Opp9yp = zero;
    if (k0 - 2 >= -1) {
        const uint32_t vp = vlay(k0 - 2);
        Edn[0] = ld(rE, vp, 0, 0);
        Edn[9] = ld(rE, vp, 9, 0);
        Edn[10] = ld(rE, vp, 10, 0);
        Edn[11] = ld(rE, vp, 11, 0);
        Eyp_dn9 = ld(rE, vp, 9, 1);
        Eym_dn10 = ld(rE, vp, 10, -1);
    }
    // rows j -+ 1 outside the box are the zero halo: their loads read zeros,
    // and the values computed for them are forced to zero
    const bool ym_dom = dom && ym_ok, yp_dom = dom && yp_ok;

    // One march step.  STEADY steps (layers k0 + 1 .. k1 - 1) have every
    // condition true at compile time, so that their body is one basic block
    // the scheduler can order freely; the two first steps and the last one
    // test them.
    auto step = [&](auto steady, int t) {
        constexpr bool S = decltype(steady)::value;
        const bool tin = S || (t >= 0 && t < L.nz);  // O'(t) is computed inside the box only
        const int te = t - 1;
        const bool ework = S || (te >= k0 && te < k1);  // E'(t-1) of this chunk
        const bool wr = S || (t >= k0 && t < k1);       // O'(t) of this chunk
        const uint32_t vt = vlay(t), vte = vt - szb, vtu = vt + szb;
        // stores of lanes that own no cell go beyond the buffer and are discarded
        const uint32_t ste = inner ? vte : 0x80000000u, sto = inner ? vt : 0x80000000u;
        const RowPot<T, POT> pe{rP, vte, srow[1], trap, dt, i, j, te};
        auto ebase = [&](int l) {
            if constexpr (kCarryE) return Edn[l];
            else return ld(rE, vte, l, 0);
        };
        auto esink = [&](int l, C v) { st(wE, ste, l, v); };

        // ---- E'(t-1), classes 1..8: no O'(t) needed ------------------------
        OffColumn<C> oo;
        oo.xm3 = shift_lane_c<true>(Op[3]);
        oo.xm4 = shift_lane_c<true>(Op[4]);
        oo.xp0 = shift_lane_c<false>(Op[0]);
        oo.xp5 = shift_lane_c<false>(Op[5]);
        oo.xp_dn0 = shift_lane_c<false>(Opp0);
        oo.ym7 = Opym7;
        oo.ym8 = Opym8;
        oo.yp1 = Opyp1;
        oo.yp9 = Opyp9;
        oo.yp_dn9 = Opp9yp;
        oo.xm_yp1 = shift_lane_c<true>(Opyp1);
        oo.xp_ym5 = shift_lane_c<false>(Opym5);
        oo.xm_up2 = zero;
        oo.ym_up6 = zero;
        if (ework) {
            update_class<1, T, FORM>(Op, Opp10, Opp11, zero, zero, oo, pe, H, ebase, esink);
            update_class<2, T, FORM>(Op, Opp10, Opp11, zero, zero, oo, pe, H, ebase, esink);
            update_class<3, T, FORM>(Op, Opp10, Opp11, zero, zero, oo, pe, H, ebase, esink);
            update_class<4, T, FORM>(Op, Opp10, Opp11, zero, zero, oo, pe, H, ebase, esink);
            update_class<5, T, FORM>(Op, Opp10, Opp11, zero, zero, oo, pe, H, ebase, esink);
            update_class<6, T, FORM>(Op, Opp10, Opp11, zero, zero, oo, pe, H, ebase, esink);
            update_class<7, T, FORM>(Op, Opp10, Opp11, zero, zero, oo, pe, H, ebase, esink);
            update_class<8, T, FORM>(Op, Opp10, Opp11, zero, zero, oo, pe, H, ebase, esink);
        }

        // ---- O'(t): own classes 2 and 6 (E'(t-1) reads them), then the
        // exported classes of rows j-1 {5..8} and j+1 {1, 9} ----------------
        C Ec[12], Oc[12];
        C nEup2 = zero, nEup6 = zero;  // E(t+1), own classes 2, 6
        C Ocym5 = zero, Ocym6 = zero, Ocym7 = zero, Ocym8 = zero, Ocyp1 = zero, Ocyp9 = zero;
        C Eym10 = zero, Eyp9 = zero;
#pragma unroll
        for (int l = 0; l < 12; l++) {
            Ec[l] = zero;
            Oc[l] = zero;
        }
        OffColumn<C> oe;
        const RowPot<T, POT> po{rP, vt, srow[1], trap, dt, i, j, t};
        auto obase = [&](int l) { return ld(rO, vt, l, 0); };
        auto osink = [&](int l, C v) {
            Oc[l] = keep_if(dom, v);
            if (wr) st(wO, sto, l, v);
        };
        if (tin) {
#pragma unroll
            for (int l = 0; l < 12; l++) Ec[l] = ld(rE, vt, l, 0);
            nEup2 = ld(rE, vtu, 2, 0);
            nEup6 = ld(rE, vtu, 6, 0);
            const C Eym5 = ld(rE, vt, 5, -1), Eym7 = ld(rE, vt, 7, -1), Eym8 = ld(rE, vt, 8, -1);
            const C nEym_up6 = ld(rE, vtu, 6, -1);
            const C Eyp1 = ld(rE, vt, 1, 1);
            Eyp9 = ld(rE, vt, 9, 1);
            oe.xm3 = shift_lane_c<true>(Ec[3]);
            oe.xm4 = shift_lane_c<true>(Ec[4]);
            oe.xp0 = shift_lane_c<false>(Ec[0]);
            oe.xp5 = shift_lane_c<false>(Ec[5]);
            oe.xm_up2 = shift_lane_c<true>(nEup2);
            oe.xp_dn0 = shift_lane_c<false>(Edn[0]);
            oe.ym7 = Eym7;
            oe.ym8 = Eym8;
            oe.yp1 = Eyp1;
            oe.yp9 = Eyp9;
            oe.ym_up6 = nEym_up6;
            oe.yp_dn9 = Eyp_dn9;
            oe.xm_yp1 = shift_lane_c<true>(Eyp1);
            oe.xp_ym5 = shift_lane_c<false>(Eym5);
            update_class<2, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<6, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            {
                // row j-1: its classes 4..8 and 10 at layer t, 10 at t-1; its
                // y+1 neighbours are this row's values (each class reads its
                // own value as well as its four neighbours)
                C Er[12];
#pragma unroll
                for (int l = 0; l < 12; l++) Er[l] = zero;
                Er[4] = ld(rE, vt, 4, -1);
                Er[5] = Eym5;
                Er[6] = ld(rE, vt, 6, -1);
                Er[7] = Eym7;
                Er[8] = Eym8;
                Eym10 = ld(rE, vt, 10, -1);
                Er[10] = Eym10;
                OffColumn<C> orr;
                orr.xm3 = orr.xp0 = orr.xp5 = orr.xm_up2 = orr.xp_dn0 = orr.ym7 = orr.ym8 = orr.ym_up6 = orr.xp_ym5 =
                    zero;
                orr.xm4 = shift_lane_c<true>(Er[4]);
                orr.xm_yp1 = shift_lane_c<true>(Ec[1]);
                orr.yp1 = Ec[1];
                orr.yp9 = Ec[9];
                orr.yp_dn9 = Edn[9];
                const RowPot<T, POT> pr{rP, vt, srow[0], trap, dt, i, j - 1, t};
                auto base = [&](int l) { return ld(rO, vt, l, -1); };
                auto sink = [&](int l, C v) {
                    v = keep_if(ym_dom, v);
                    if (l == 5) Ocym5 = v;
                    else if (l == 6) Ocym6 = v;
                    else if (l == 7) Ocym7 = v;
                    else Ocym8 = v;
                };
                update_class<5, T, FORM>(Er, Eym_dn10, zero, zero, zero, orr, pr, H, base, sink);
                update_class<6, T, FORM>(Er, Eym_dn10, zero, zero, zero, orr, pr, H, base, sink);
                update_class<7, T, FORM>(Er, Eym_dn10, zero, zero, zero, orr, pr, H, base, sink);
                update_class<8, T, FORM>(Er, Eym_dn10, zero, zero, zero, orr, pr, H, base, sink);
            }
            {
                // row j+1: its classes 0..3, 9 and 11 at layer t; its y-1
                // neighbours are this row's values
                C Er[12];
#pragma unroll
                for (int l = 0; l < 12; l++) Er[l] = zero;
                Er[0] = ld(rE, vt, 0, 1);
                Er[1] = Eyp1;
                Er[2] = ld(rE, vt, 2, 1);
                Er[3] = ld(rE, vt, 3, 1);
                Er[9] = Eyp9;
                Er[11] = ld(rE, vt, 11, 1);
                OffColumn<C> orr;
                orr.xm3 = orr.xm4 = orr.xp0 = orr.xp5 = orr.xm_up2 = orr.xp_dn0 = orr.yp1 = orr.yp9 = orr.yp_dn9 =
                    orr.xm_yp1 = zero;
                orr.xp_ym5 = shift_lane_c<false>(Ec[5]);
                orr.ym8 = Ec[8];
                orr.ym7 = Ec[7];
                orr.ym_up6 = nEup6;
                const RowPot<T, POT> pr{rP, vt, srow[2], trap, dt, i, j + 1, t};
                auto base = [&](int l) { return ld(rO, vt, l, 1); };
                auto sink = [&](int l, C v) {
                    v = keep_if(yp_dom, v);
                    if (l == 1) Ocyp1 = v;
                    else Ocyp9 = v;
                };
                update_class<1, T, FORM>(Er, zero, zero, zero, zero, orr, pr, H, base, sink);
                update_class<9, T, FORM>(Er, zero, zero, zero, zero, orr, pr, H, base, sink);
            }
        }

        // ---- E'(t-1), classes 0, 9, 10, 11: these read O'(t) -----------------
        if (ework) {
            oo.xm_up2 = shift_lane_c<true>(Oc[2]);
            oo.ym_up6 = Ocym6;
            update_class<0, T, FORM>(Op, Opp10, Opp11, Oc[2], Oc[6], oo, pe, H, ebase, esink);
            update_class<9, T, FORM>(Op, Opp10, Opp11, Oc[2], Oc[6], oo, pe, H, ebase, esink);
            update_class<10, T, FORM>(Op, Opp10, Opp11, Oc[2], Oc[6], oo, pe, H, ebase, esink);
            update_class<11, T, FORM>(Op, Opp10, Opp11, Oc[2], Oc[6], oo, pe, H, ebase, esink);
        }
        // O'(t-1) is no longer read but for the classes carried to t+1
        Opp0 = Op[0];
        Opp10 = Op[10];
        Opp11 = Op[11];
        Opp9yp = Opyp9;

        // ---- O'(t): the other own classes -----------------------------------
        if (tin) {
            update_class<0, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<1, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<3, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<4, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<5, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<7, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<8, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<9, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<10, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
            update_class<11, T, FORM>(Ec, Edn[10], Edn[11], nEup2, nEup6, oe, po, H, obase, osink);
        }

        // ---- advance --------------------------------------------------------
#pragma unroll
        for (int l = 0; l < 12; l++) Op[l] = Oc[l];
        Opym5 = Ocym5;
        Opym7 = Ocym7;
        Opym8 = Ocym8;
        Opyp1 = Ocyp1;
        Opyp9 = Ocyp9;
#pragma unroll
        for (int l = 0; l < 12; l++)
            if (kCarryE || l == 0 || l == 9 || l == 10 || l == 11) Edn[l] = Ec[l];
        Eyp_dn9 = Eyp9;
        Eym_dn10 = Eym10;
    };

    using Edge = std::integral_constant<bool, false>;
    using Steady = std::integral_constant<bool, true>;
    int t = k0 - 1;
    for (; t <= k0; t++) step(Edge{}, t);
    for (; t < k1; t++) step(Steady{}, t);
    step(Edge{}, k1);
}

}  // namespace gpe
