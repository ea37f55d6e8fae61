// gpe_fused.hpp -- one z-march pass that performs a whole leapfrog step:
//   O' = O - i dt H(E) E        (odd half-update, GPUGrossPitaevskii_BCC.cu:478)
//   E' = E - i dt H(O') O'      (even half-update, GPUGrossPitaevskii_BCC.cu:480)
// reading E and O once and writing E' and O' once (out of place), instead of
// two passes that each read two fields and write one.  Per node-step that is
// 4 field transfers instead of 6 (64 B instead of 96 B in FP64).
//
// Tiling.  A block of 64 x TY threads owns a footprint of (w+2) x (h+2)
// columns: an interior of w x h columns whose E' it writes, plus a one-column
// ring.  It marches the layers of a z-chunk; at step t it computes O'(t) on
// the whole footprint and E'(t-1) on the interior.  The ring's O' is
// recomputed redundantly by each block that needs it.  On this mesh every
// O' class a ring column exports to the interior depends only on data on
// the interior side of that column (e.g. the classes x+1 reads from x are
// {0, 5} at layer t-1 and {0} at t-2, whose four faces never point to x+1),
// so ring values need nothing beyond the footprint.
//
// Exchange.  x-neighbours are adjacent lanes of one wavefront: DPP wave
// shifts (v_mov_b32_dpp wave_shr:1 / wave_shl:1), no memory traffic.
// y-neighbours of E come from global loads (L1/L2 hits: the neighbouring
// wave just read those rows); y-neighbours of O' go through LDS, classes
// {1,5,6,7,8,9} of two layers.  z-neighbours stay in registers.
#pragma once
#include "gpe_kernels.hpp"

namespace gpe {

// 128-bit buffer store.  A store of more than 64 bits reads its data
// registers after it issues, so an instruction that writes them right after
// the store needs a wait state in between.  The compiler inserts it only when
// the store's scalar offset is the constant 0 (with a scalar register there,
// the next instruction may overwrite the data in flight and the store writes
// new values for some lanes), so the scalar offset goes into the per-lane one.
template <int AUX = 0, typename V>
__device__ __forceinline__ void buffer_store_128(V data, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    static_assert(sizeof(V) == 16, "128-bit store");
    using W = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(W, data), r, voff + soff, 0, AUX);
}

// lane i <- lane i-1 (value of the x-1 neighbour); lane 0 gets 0 (bound
// control: a lane whose source is outside the wave reads zero, so no old
// value has to be set first).
__device__ __forceinline__ int dpp_from_xm(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
// lane i <- lane i+1 (value of the x+1 neighbour); lane 63 gets 0.
__device__ __forceinline__ int dpp_from_xp(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }

template <bool FROM_XM>
__device__ __forceinline__ float shift_lane(float v) {
    const int i = __float_as_int(v);
    return __int_as_float(FROM_XM ? dpp_from_xm(i) : dpp_from_xp(i));
}
template <bool FROM_XM>
__device__ __forceinline__ double shift_lane(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
    const int rlo = FROM_XM ? dpp_from_xm(lo) : dpp_from_xp(lo);
    const int rhi = FROM_XM ? dpp_from_xm(hi) : dpp_from_xp(hi);
    return __longlong_as_double(((long long)rhi << 32) | (unsigned int)rlo);
}
template <bool FROM_XM, typename C>
__device__ __forceinline__ C shift_lane_c(C v) {
    C r;
    r.x = shift_lane<FROM_XM>(v.x);
    r.y = shift_lane<FROM_XM>(v.y);
    return r;
}

// The 14 off-column neighbour values one half-update of a cell needs, named
// by their (dx,dy,dz) offset and class (see bcc_face).
template <typename C>
struct OffColumn {
    C xm3, xm4, xp0, xp5;  // (-1,0,0)->3,4   (+1,0,0)->0,5
    C xm_up2, xp_dn0;      // (-1,0,+1)->2    (+1,0,-1)->0
    C ym7, ym8, yp1, yp9;  // (0,-1,0)->7,8   (0,+1,0)->1,9
    C ym_up6, yp_dn9;      // (0,-1,+1)->6    (0,+1,-1)->9
    C xm_yp1, xp_ym5;      // (-1,+1,0)->1    (+1,-1,0)->5
};

// Value of face f of class l given the own column (cur = layer c, dn = layer
// c-1 classes 0/10/11, up = layer c+1 classes 2/6) and the off-column set.
template <int L, int F, typename C>
__device__ __forceinline__ C face_value(const C (&cur)[12], const C &dn10, const C &dn11, const C &up2, const C &up6,
                                        const OffColumn<C> &o) {
    constexpr Face e = bcc_face(L, F);
    if constexpr (e.dx == 0 && e.dy == 0 && e.dz == 0) return cur[e.m];
    else if constexpr (e.dx == 0 && e.dy == 0 && e.dz == -1) return e.m == 10 ? dn10 : dn11;
    else if constexpr (e.dx == 0 && e.dy == 0 && e.dz == 1) return e.m == 2 ? up2 : up6;
    else if constexpr (e.dx == -1 && e.dy == 0 && e.dz == 0) return e.m == 3 ? o.xm3 : o.xm4;
    else if constexpr (e.dx == 1 && e.dy == 0 && e.dz == 0) return e.m == 0 ? o.xp0 : o.xp5;
    else if constexpr (e.dx == -1 && e.dz == 1) return o.xm_up2;
    else if constexpr (e.dx == 1 && e.dz == -1) return o.xp_dn0;
    else if constexpr (e.dx == 0 && e.dy == -1 && e.dz == 0) return e.m == 7 ? o.ym7 : o.ym8;
    else if constexpr (e.dx == 0 && e.dy == 1 && e.dz == 0) return e.m == 1 ? o.yp1 : o.yp9;
    else if constexpr (e.dy == -1 && e.dz == 1) return o.ym_up6;
    else if constexpr (e.dy == 1 && e.dz == -1) return o.yp_dn9;
    else if constexpr (e.dx == -1 && e.dy == 1) return o.xm_yp1;
    else return o.xp_ym5;
}

// V*dt of the 12 classes of one cell: in-kernel trap or streamed array.
template <typename T, int POT>
struct PotCell {
    const T *__restrict__ pot;
    long long base, xp;
    const Trap &trap;
    double dt;
    int i, j, k;
    __device__ __forceinline__ T operator[](int l) const {
        if (POT == POT_TRAP) return (T)(trap_v(trap, i, j, k, l) * dt);
        return pot[base + l * xp];
    }
};

// One real-time half-update of class L of one cell: sink(L, base(L) - i*s)
// with s = dt*H(psi)psi, psi = cur[L], neighbours from the own column and
// the off-column set.  base and sink are per-class callables so that the
// read-modify-write operand is loaded right before use and the result
// leaves registers right after.
template <int L, typename T, int FORM, typename P, typename B, typename S>
__device__ __forceinline__ void update_class(const cplx_t<T> (&cur)[12], const cplx_t<T> &dn10,
                                             const cplx_t<T> &dn11, const cplx_t<T> &up2, const cplx_t<T> &up6,
                                             const OffColumn<cplx_t<T>> &o, const P &pot,
                                             const Hamiltonian<T, FORM> &H, B &&base, S &&sink) {
#pragma clang fp contract(off)
    cplx_t<T> n[4];
    n[0] = face_value<L, 0>(cur, dn10, dn11, up2, up6, o);
    n[1] = face_value<L, 1>(cur, dn10, dn11, up2, up6, o);
    n[2] = face_value<L, 2>(cur, dn10, dn11, up2, up6, o);
    n[3] = face_value<L, 3>(cur, dn10, dn11, up2, up6, o);
    const cplx_t<T> s = H.apply(n, cur[L], pot[L]);
    const cplx_t<T> b = base(L);
    cplx_t<T> r;
    r.x = b.x + s.y;
    r.y = b.y + (-s.x);
    sink(L, r);
}

template <typename T, int FORM, typename P, typename B, typename S>
__device__ __forceinline__ void update_all(const cplx_t<T> (&cur)[12], const cplx_t<T> &dn10, const cplx_t<T> &dn11,
                                           const cplx_t<T> &up2, const cplx_t<T> &up6,
                                           const OffColumn<cplx_t<T>> &o, const P &pot,
                                           const Hamiltonian<T, FORM> &H, B &&base, S &&sink) {
    update_class<0>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<1>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<2>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<3>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<4>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<5>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<6>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<7>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<8>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<9>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<10>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
    update_class<11>(cur, dn10, dn11, up2, up6, o, pot, H, base, sink);
}

// Tile geometry of the fused pass (host computes it, see make_fused_grid).
struct FusedGrid {
    int ntx, nty, ntz;  // tiles per axis
    int w, h, cz;       // interior width (x), height (y), layers (z) per tile
    int remap;          // XCD-contiguous block order
};

// LDS slot of the y-exported O' classes {1,5,6,7,8,9}.
__host__ __device__ constexpr int yslot(int m) {
    return m == 1 ? 0 : m == 5 ? 1 : m == 6 ? 2 : m == 7 ? 3 : m == 8 ? 4 : 5;
}

// Logical tile index -> (bx, by, bz), with the XCD-contiguous order of
// block_coords when remap is set.
__device__ __forceinline__ int remap_block(int lin, int nb, int remap) {
    if (!remap) return lin;
    const int q = nb >> 3, r = nb & 7, xcd = lin & 7, slot = lin >> 3;
    return (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// One tile of the fused step (the whole block works on tile `w`).
template <typename T, int FORM, int POT, int TY>
__device__ __forceinline__ void fused_tile(cplx_t<T> *__restrict__ eout, cplx_t<T> *__restrict__ oout,
                                           const cplx_t<T> *__restrict__ ein, const cplx_t<T> *__restrict__ oin,
                                           const T *__restrict__ pot, const Layout &L, const Hamiltonian<T, FORM> &H,
                                           const Trap &trap, double dt, const FusedGrid &G, int w,
                                           cplx_t<T> (&ylds)[2][6][TY][64]) {
#pragma clang fp contract(off)
    using C = cplx_t<T>;
    const int bx = w % G.ntx, by = (w / G.ntx) % G.nty, bz = w / (G.ntx * G.nty);
    const int xs = bx * G.w, ys = by * G.h, k0 = bz * G.cz;
    const int wx = min(G.w, L.nx - xs), hy = min(G.h, L.ny - ys);
    const int k1 = min(k0 + G.cz, L.nz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = xs - 1 + tx, j = ys - 1 + ty;
    const bool foot = tx <= wx + 1 && ty <= hy + 1;
    const bool dom = foot && i >= 0 && i < L.nx && j >= 0 && j < L.ny;
    const bool inner = dom && tx >= 1 && tx <= wx && ty >= 1 && ty <= hy;
    const long long col = L.origin + (long long)j * L.sy + i;  // valid when dom

    const C zero = {(T)0, (T)0};
    auto ld = [&](const C *__restrict__ f, int k, int l, long long off) -> C {
        return f[col + (long long)k * L.sz + l * L.xp + off];
    };

    // ---- march state ------------------------------------------------------
    C Ec[12];               // E(t)
    C Ep0 = zero, Ep10 = zero, Ep11 = zero;  // E(t-1) classes 0, 10, 11
    C Eup2 = zero, Eup6 = zero;   // E(t+1) classes 2, 6
    C Oc[12], Op[12];       // O'(t), O'(t-1)
    C Opp0 = zero, Opp10 = zero, Opp11 = zero;  // O'(t-2) classes 0, 10, 11
    C E_yp9_prev = zero;    // E(t-1) class 9 of the y+1 column
    C O_yp9_prev = zero;    // O'(t-2) class 9 of the y+1 column
#pragma unroll
    for (int l = 0; l < 12; l++) {
        Ec[l] = zero;
        Oc[l] = zero;
        Op[l] = zero;
    }
    // Prologue: E(k0-2) into Ep (only classes 0,10,11 and the y+1 class 9
    // are read from it), E(k0-1) 2/6 lookahead.
    {
        const int kp = k0 - 2, kc = k0 - 1;
        if (dom && kp >= 0) {
            Ep0 = ld(ein, kp, 0, 0);
            Ep10 = ld(ein, kp, 10, 0);
            Ep11 = ld(ein, kp, 11, 0);
            E_yp9_prev = ld(ein, kp, 9, L.sy);
        }
        if (dom && kc >= 0) {
            Eup2 = ld(ein, kc, 2, 0);
            Eup6 = ld(ein, kc, 6, 0);
        }
    }

    for (int t = k0 - 1; t <= k1; t++) {
        const bool tin = t >= 0 && t < L.nz;
        // ---- loads of layer t (E rest, O all) and lookahead E(t+1){2,6} ----
        OffColumn<C> oe;
        C E_yp9_now = zero;
        oe.ym7 = oe.ym8 = oe.yp1 = oe.yp9 = oe.ym_up6 = oe.xm_yp1 = oe.xp_ym5 = zero;
        C nEup2 = zero, nEup6 = zero;
        // lookahead E(t+1) classes 2 and 6: needed also when t itself is the
        // zero halo layer below the box (t = -1), for the next step's E(t)
        if (dom && t + 1 >= 0 && t + 1 < L.nz) {
            nEup2 = ld(ein, t + 1, 2, 0);
            nEup6 = ld(ein, t + 1, 6, 0);
            if (tin) oe.ym_up6 = ld(ein, t + 1, 6, -L.sy);
        }
        if (dom && tin) {
#pragma unroll
            for (int l = 0; l < 12; l++) {
                if (l == 2) Ec[l] = Eup2;
                else if (l == 6) Ec[l] = Eup6;
                else Ec[l] = ld(ein, t, l, 0);
            }
            oe.ym7 = ld(ein, t, 7, -L.sy);
            oe.ym8 = ld(ein, t, 8, -L.sy);
            oe.yp1 = ld(ein, t, 1, L.sy);
            E_yp9_now = ld(ein, t, 9, L.sy);
            oe.xm_yp1 = ld(ein, t, 1, L.sy - 1);
            oe.xp_ym5 = ld(ein, t, 5, 1 - L.sy);
        } else {
#pragma unroll
            for (int l = 0; l < 12; l++) Ec[l] = zero;
        }
        oe.yp9 = E_yp9_now;
        oe.yp_dn9 = E_yp9_prev;
        // x-neighbours of E by lane shifts (all lanes execute these)
        oe.xm3 = shift_lane_c<true>(Ec[3]);
        oe.xm4 = shift_lane_c<true>(Ec[4]);
        oe.xp0 = shift_lane_c<false>(Ec[0]);
        oe.xp5 = shift_lane_c<false>(Ec[5]);
        oe.xm_up2 = shift_lane_c<true>(nEup2);
        oe.xp_dn0 = shift_lane_c<false>(Ep0);

        // ---- O'(t) on the footprint -------------------------------------
        if (dom && tin) {
            const PotCell<T, POT> pv{pot, col + (long long)t * L.sz, L.xp, trap, dt, i, j, t};
            const bool wr = inner && t >= k0 && t < k1;
            const long long at = col + (long long)t * L.sz;
            update_all<T, FORM>(
                Ec, Ep10, Ep11, nEup2, nEup6, oe, pv, H, [&](int l) { return oin[at + l * L.xp]; },
                [&](int l, C v) {
                    Oc[l] = v;
                    if (wr) oout[at + l * L.xp] = v;
                });
        } else {
#pragma unroll
            for (int l = 0; l < 12; l++) Oc[l] = zero;
        }
        // publish O'(t) classes read by the y-1 / y+1 rows
        {
            const int s = t & 1;
            ylds[s][yslot(1)][ty][tx] = Oc[1];
            ylds[s][yslot(5)][ty][tx] = Oc[5];
            ylds[s][yslot(6)][ty][tx] = Oc[6];
            ylds[s][yslot(7)][ty][tx] = Oc[7];
            ylds[s][yslot(8)][ty][tx] = Oc[8];
            ylds[s][yslot(9)][ty][tx] = Oc[9];
        }
        __syncthreads();

        // ---- E'(t-1) on the interior -------------------------------------
        OffColumn<C> oo;
        oo.xm3 = shift_lane_c<true>(Op[3]);
        oo.xm4 = shift_lane_c<true>(Op[4]);
        oo.xp0 = shift_lane_c<false>(Op[0]);
        oo.xp5 = shift_lane_c<false>(Op[5]);
        oo.xm_up2 = shift_lane_c<true>(Oc[2]);
        oo.xp_dn0 = shift_lane_c<false>(Opp0);
        C O_yp9_now = zero;
        {
            const int sp = (t - 1) & 1, sc = t & 1;
            const int tym = ty > 0 ? ty - 1 : 0, typ = ty < TY - 1 ? ty + 1 : TY - 1;
            const int txm = tx > 0 ? tx - 1 : 0, txp = tx < 63 ? tx + 1 : 63;
            oo.ym7 = ylds[sp][yslot(7)][tym][tx];
            oo.ym8 = ylds[sp][yslot(8)][tym][tx];
            oo.ym_up6 = ylds[sc][yslot(6)][tym][tx];
            oo.yp1 = ylds[sp][yslot(1)][typ][tx];
            O_yp9_now = ylds[sp][yslot(9)][typ][tx];
            oo.xm_yp1 = ylds[sp][yslot(1)][typ][txm];
            oo.xp_ym5 = ylds[sp][yslot(5)][tym][txp];
        }
        oo.yp9 = O_yp9_now;
        oo.yp_dn9 = O_yp9_prev;
        const int te = t - 1;
        if (inner && te >= k0 && te < k1) {
            const PotCell<T, POT> pv{pot, col + (long long)te * L.sz, L.xp, trap, dt, i, j, te};
            const long long at = col + (long long)te * L.sz;
            update_all<T, FORM>(
                Op, Opp10, Opp11, Oc[2], Oc[6], oo, pv, H, [&](int l) { return ein[at + l * L.xp]; },
                [&](int l, C v) { eout[at + l * L.xp] = v; });
        }
        __syncthreads();  // the next step overwrites the slot read above

        // ---- advance -----------------------------------------------------
        Opp0 = Op[0];
        Opp10 = Op[10];
        Opp11 = Op[11];
        Ep0 = Ec[0];
        Ep10 = Ec[10];
        Ep11 = Ec[11];
#pragma unroll
        for (int l = 0; l < 12; l++) Op[l] = Oc[l];
        Eup2 = nEup2;
        Eup6 = nEup6;
        E_yp9_prev = E_yp9_now;
        O_yp9_prev = O_yp9_now;
    }
}

template <typename T, int FORM, int POT, int TY>
__global__ void __launch_bounds__(64 * TY)
    k_fused(cplx_t<T> *__restrict__ eout, cplx_t<T> *__restrict__ oout, const cplx_t<T> *__restrict__ ein,
            const cplx_t<T> *__restrict__ oin, const T *__restrict__ pot, Layout L, Hamiltonian<T, FORM> H, Trap trap,
            double dt, FusedGrid G) {
    __shared__ cplx_t<T> ylds[2][6][TY][64];
    const int w = remap_block(blockIdx.x, G.ntx * G.nty * G.ntz, G.remap);
    fused_tile<T, FORM, POT, TY>(eout, oout, ein, oin, pot, L, H, trap, dt, G, w, ylds);
}

}  // namespace gpe
