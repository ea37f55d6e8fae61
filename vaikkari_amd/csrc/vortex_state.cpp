// vortex_state.cpp -- generating a VortexState: the start shape, resampling
// from a coarser state, the relaxation of the stationary axisymmetric
// Gross-Pitaevskii equation and its residual, the binary writer, and the
// two-stage recipe that produces the state.dat the stepper starts from
// (koodit/GrossPitaevskii/VortexState.cpp:5-47, 74-196, 250-257, 282-311,
// 342-361; GPUGrossPitaevskii_BCC.cu:516-535).
//
// The table is f(r_i, z_j), r_i = rplug + i*rstep, z_j = j*zstep, stored
// r-major (index i*zsize + j).  The relaxation solves
//   (-0.5 lap_rz + kappa^2/(2 r^2) + V(r,z) + g f^2) f = mu f
// with second-order differences in r and z, f = 0 on the r = rplug row and
// beyond the table, and the mirror f(r,-dz) = f(r,+dz) at z = 0.
//
// Each update of a sweep reads the values the same sweep has just produced
// at (r - dr, z) and (r, z - dz) (Gauss-Seidel), and mu and the
// normalisation of every sweep come from sums whose order fixes the last
// bits of the result, so the sweep order is part of the algorithm.  The
// table holds a few hundred thousand doubles (cache resident) and stays on
// the host; the 3-D field it seeds is what runs on the GPU.  Large tables
// are relaxed by a team of threads without changing a bit of the result:
// rows go round-robin to the threads, and row i updates column j only after
// row i-1 has passed it (a pipelined wavefront), so every update reads the
// same values as in the serial sweep; the terms of the three sums are
// computed in parallel and added by one thread in the serial order.
// Compiled with -ffp-contract=off, every expression below is evaluated in
// the grouping written, which is the reference's.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <thread>
#include <vector>

#include "gpe_internal.hpp"
#include "vortex_state.hpp"

namespace gpe {
namespace {

constexpr double kPi = 3.1415926535897932385;
constexpr double kTwoPi = 6.2831853071795864770;

// V(r, z) of the relaxation: a caller's function, or the trap
// 0.5(r^2 + ratio^2 z^2) of potentialRZ (GPUGrossPitaevskii_BCC.cu:28-31).
struct RzPotential {
    gpe_rz_potential fn;
    double ratio;
    void *user;
    double operator()(double r, double z) const {
        if (fn) return fn(r, z, user);
        return 0.5 * (r * r + ratio * ratio * z * z);
    }
};

// Finite-difference coefficients of the operator on radius row i; the
// kinetic prefactor `half` is 0.5 for H (relaxation, residual) and 1 for
// the curvature.  dv is the ring volume 2 pi r dr dz of one cell; the
// z = 0 cell counts once, every other z cell twice (z and -z).
struct RzRow {
    double r, diag, grad, dv;
};

struct RzOperator {
    const VortexState &v;
    double cr, cz, kk, w0;
    explicit RzOperator(const VortexState &vs)
        : v(vs),
          cr(0.5 / (vs.rstep * vs.rstep)),
          cz(vs.zsize > 1 ? 0.5 / (vs.zstep * vs.zstep) : 0.0),
          kk(0.5 * (double)(vs.kappa * vs.kappa)),
          w0(kTwoPi * vs.rstep * vs.zstep) {}
    RzRow row(uint32_t i) const {
        RzRow q;
        q.r = v.rplug + i * v.rstep;
        q.diag = kk / (q.r * q.r) + 2.0 * (cr + cz);
        q.grad = 0.25 / (q.r * v.rstep);
        q.dv = w0 * q.r;
        return q;
    }
};

// The four neighbours of table entry (i, j), with the boundary rules above.
struct RzNeighbours {
    double dn_r, up_r, dn_z, up_z;
};

inline RzNeighbours neighbours(const VortexState &v, uint32_t i, uint32_t j, size_t ii) {
    RzNeighbours n;
    n.up_r = (i + 1 < v.rsize ? v.f[ii + v.zsize] : 0.0);
    n.dn_r = v.f[ii - v.zsize];
    n.up_z = (j + 1 < v.zsize ? v.f[ii + 1] : 0.0);
    n.dn_z = (j > 0 ? v.f[ii - 1] : n.up_z);
    return n;
}

// Off-diagonal part of H f at (i, j) plus `a`, grouped as the reference
// writes a + grad(dn_r - up_r) - cr(dn_r + up_r) - cz(dn_z + up_z).
inline double off_plus(double a, const RzOperator &op, const RzRow &q, const RzNeighbours &n) {
    return ((a + q.grad * (n.dn_r - n.up_r)) - op.cr * (n.dn_r + n.up_r)) - op.cz * (n.dn_z + n.up_z);
}

// One Gauss-Seidel sweep (VortexState.cpp:123-145): solves each entry's row
// of (H - mu) f = 0 for f(i,j) with the nonlinearity frozen at the old value.
// Returns sum (f_old - f_new)^2 dv in table order.
double relax_sweep(VortexState &v, const RzOperator &op, const RzPotential &V) {
    double diff = 0.0;
    size_t ii = v.zsize;
    for (uint32_t i = 1; i < v.rsize; i++) {
        const RzRow q = op.row(i);
        double dv = q.dv;
        for (uint32_t j = 0; j < v.zsize; j++, ii++) {
            const double pot = V(q.r, j * v.zstep) + v.g * v.f[ii] * v.f[ii];
            const RzNeighbours n = neighbours(v, i, j, ii);
            const double num = (q.grad * (n.dn_r - n.up_r) - op.cr * (n.dn_r + n.up_r)) - op.cz * (n.dn_z + n.up_z);
            const double fresh = num / (v.mu - q.diag - pot);
            const double d = v.f[ii] - fresh;
            v.f[ii] = fresh;
            diff += d * d * dv;
            if (j == 0) dv *= 2.0;
        }
    }
    return diff;
}

// <f, f> and <f, H f> in table order (VortexState.cpp:147-168).
void rayleigh_sums(const VortexState &v, const RzOperator &op, const RzPotential &V, double &prob, double &prob2) {
    prob = 0.0;
    prob2 = 0.0;
    size_t ii = v.zsize;
    for (uint32_t i = 1; i < v.rsize; i++) {
        const RzRow q = op.row(i);
        double dv = q.dv;
        for (uint32_t j = 0; j < v.zsize; j++, ii++) {
            const double fij = v.f[ii];
            const double pot = V(q.r, j * v.zstep) + v.g * fij * fij;
            const RzNeighbours n = neighbours(v, i, j, ii);
            prob += fij * fij * dv;
            prob2 += off_plus((pot + q.diag) * fij, op, q, n) * fij * dv;
            if (j == 0) dv *= 2.0;
        }
    }
}

// Relaxation by a team of threads, bit-identical to the serial loop in
// gpe_vortex_state_iterate (see the file comment).  The calling thread is
// worker 0; it sums, calls `progress` and decides when to stop.
struct RelaxTeam {
    VortexState &v;
    const RzOperator &op;
    const RzPotential &V;
    const int T;                 // workers
    static constexpr int kChunk = 32;  // columns between progress updates
    std::unique_ptr<std::atomic<int>[]> done;  // per row: columns finished this sweep
    std::vector<double> dd, t1, t2;             // per entry: sweep change, <f,f> and <f,Hf> terms
    std::atomic<int> arrived{0}, phase{0};
    std::atomic<int> stop{0};
    double diff = 0.0, prob = 0.0, prob2 = 0.0;

    RelaxTeam(VortexState &vs, const RzOperator &o, const RzPotential &pot, int workers)
        : v(vs), op(o), V(pot), T(workers), done(new std::atomic<int>[vs.rsize]),
          dd(vs.f.size()), t1(vs.f.size()), t2(vs.f.size()) {
        for (uint32_t i = 0; i < v.rsize; i++) done[i].store(i == 0 ? (int)v.zsize : 0);
    }

    // Sense-reversing barrier over the T workers.
    void barrier() {
        const int ph = phase.load(std::memory_order_acquire);
        if (arrived.fetch_add(1, std::memory_order_acq_rel) == T - 1) {
            arrived.store(0, std::memory_order_relaxed);
            phase.store(ph + 1, std::memory_order_release);
        } else {
            int spins = 0;
            while (phase.load(std::memory_order_acquire) == ph)
                if (++spins > 64) std::this_thread::yield();
        }
    }

    void sweep_rows(int t) {
        for (uint32_t i = 1 + t; i < v.rsize; i += T) {
            const RzRow q = op.row(i);
            const size_t base = (size_t)i * v.zsize;
            for (int j0 = 0; j0 < (int)v.zsize; j0 += kChunk) {
                const int j1 = std::min((int)v.zsize, j0 + kChunk);
                int spins = 0;
                while (done[i - 1].load(std::memory_order_acquire) < j1)
                    if (++spins > 64) std::this_thread::yield();
                for (int j = j0; j < j1; j++) {
                    const size_t ii = base + j;
                    const double pot = V(q.r, j * v.zstep) + v.g * v.f[ii] * v.f[ii];
                    const RzNeighbours n = neighbours(v, i, (uint32_t)j, ii);
                    const double num =
                        (q.grad * (n.dn_r - n.up_r) - op.cr * (n.dn_r + n.up_r)) - op.cz * (n.dn_z + n.up_z);
                    const double fresh = num / (v.mu - q.diag - pot);
                    const double d = v.f[ii] - fresh;
                    v.f[ii] = fresh;
                    dd[ii] = d * d * (j == 0 ? q.dv : 2.0 * q.dv);
                }
                done[i].store(j1, std::memory_order_release);
            }
        }
    }

    void term_rows(int t) {
        for (uint32_t i = 1 + t; i < v.rsize; i += T) {
            const RzRow q = op.row(i);
            for (uint32_t j = 0; j < v.zsize; j++) {
                const size_t ii = (size_t)i * v.zsize + j;
                const double dv = j == 0 ? q.dv : 2.0 * q.dv;
                const double fij = v.f[ii];
                const double pot = V(q.r, j * v.zstep) + v.g * fij * fij;
                const RzNeighbours n = neighbours(v, i, j, ii);
                t1[ii] = fij * fij * dv;
                t2[ii] = off_plus((pot + q.diag) * fij, op, q, n) * fij * dv;
            }
        }
    }

    void scale_rows(int t, double scale) {
        for (uint32_t i = 1 + t; i < v.rsize; i += T) {
            double *row = v.f.data() + (size_t)i * v.zsize;
            for (uint32_t j = 0; j < v.zsize; j++) row[j] *= scale;
            done[i].store(0, std::memory_order_relaxed);
        }
    }

    // Returns the iterations run (the last one converged when stop == 2).
    uint32_t run(int t, uint32_t maxiters, double eps, gpe_relax_progress progress, void *puser) {
        uint32_t it = 0;
        for (it = 0; it < maxiters; it++) {
            sweep_rows(t);
            barrier();
            term_rows(t);
            barrier();
            if (t == 0) {
                double a = 0.0, b = 0.0, c = 0.0;
                for (size_t ii = v.zsize; ii < v.f.size(); ii++) {
                    a += dd[ii];
                    b += t1[ii];
                    c += t2[ii];
                }
                diff = a;
                prob = b;
                prob2 = c;
                v.mu = prob2 / prob;
                if (progress && it % 1000 == 0) progress(it, diff, prob, puser);
                stop.store(diff < eps ? 2 : 0, std::memory_order_relaxed);
            }
            barrier();
            scale_rows(t, 1.0 / std::sqrt(prob));
            barrier();
            if (stop.load(std::memory_order_relaxed)) return it + 1;
        }
        return it;
    }
};

// Workers for a table of this size: one per hardware thread (at most 16 and
// one per 4 rows); small tables stay serial.
int relax_workers(const VortexState &v, uint32_t requested) {
    if (requested) return (int)std::max<uint32_t>(1, std::min<uint32_t>(requested, std::max<uint32_t>(1, v.rsize - 1)));
    if ((size_t)v.rsize * v.zsize < 40000) return 1;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    return (int)std::max<uint32_t>(1, std::min<uint32_t>({hw, 16u, (v.rsize - 1) / 4}));
}

// || (H - mu) f ||^2 (VortexState.cpp:282-311).
double residual(const VortexState &v, const RzPotential &V) {
    const RzOperator op(v);
    double err = 0.0;
    size_t ii = v.zsize;
    for (uint32_t i = 1; i < v.rsize; i++) {
        const RzRow q = op.row(i);
        double dv = q.dv;
        for (uint32_t j = 0; j < v.zsize; j++, ii++) {
            const double fij = v.f[ii];
            const double pot = V(q.r, j * v.zstep) + v.g * fij * fij;
            const RzNeighbours n = neighbours(v, i, j, ii);
            const double df = off_plus((pot + q.diag - v.mu) * fij, op, q, n);
            err += df * df * dv;
            if (j == 0) dv *= 2.0;
        }
    }
    return err;
}

// setDimensions (VortexState.cpp:12-40): clamps the sizes, then the start
// shape (1 - cos(2 pi i/rsize)) (1 + cos(pi j/zsize)), normalised with the
// same ring volumes (the z = 0 row once), and mu = 1.
void set_dimensions(VortexState &v, uint32_t rsize, double rstep, double rplug, uint32_t zsize, double zstep) {
    v.rsize = rsize > 1 ? rsize : 1;
    v.rstep = rstep > 1e-13 ? rstep : 1e-13;
    v.rplug = rplug > 0.0 ? rplug : 0.0;
    v.zsize = zsize > 1 ? zsize : 1;
    v.zstep = zstep > 1e-13 ? zstep : 1e-13;
    v.mu = 1.0;
    v.f.assign((size_t)v.rsize * v.zsize, 0.0);
    double prob = 0.0;
    size_t ii = 0;
    for (uint32_t i = 0; i < v.rsize; i++) {
        const double shape_r = 1.0 - std::cos(i * kTwoPi / (double)v.rsize);
        const double r = v.rplug + i * v.rstep;
        uint32_t times = 1;
        for (uint32_t j = 0; j < v.zsize; j++, ii++) {
            v.f[ii] = (1.0 + std::cos(j * kPi / (double)v.zsize)) * shape_r;
            prob += times * v.f[ii] * v.f[ii] * r;
            times = 2;
        }
    }
    prob *= kTwoPi * v.rstep * v.zstep;
    const double scale = 1.0 / std::sqrt(prob);
    for (double &x : v.f) x *= scale;
}

}  // namespace
}  // namespace gpe

using namespace gpe;

extern "C" {

// VortexState() (VortexState.cpp:5-10): kappa 2, g 0, a 100 x 1 table of
// step 0.1 from r = 0.
gpe_status gpe_vortex_state_create(gpe_vortex_state **out) {
    if (!out) {
        set_error("gpe_vortex_state_create: null argument");
        return GPE_EINVAL;
    }
    auto *h = new gpe_vortex_state();
    h->vs.kappa = 2;
    h->vs.g = 0.0;
    set_dimensions(h->vs, 100, 0.1, 0.0, 1, 1.0);
    *out = h;
    return GPE_OK;
}

void gpe_vortex_state_set_kappa(gpe_vortex_state *h, uint32_t kappa) { h->vs.kappa = kappa; }
void gpe_vortex_state_set_threads(gpe_vortex_state *h, uint32_t threads) { h->threads = threads; }
void gpe_vortex_state_set_g(gpe_vortex_state *h, double g) { h->vs.g = g; }

gpe_status gpe_vortex_state_set_dimensions(gpe_vortex_state *h, uint32_t rsize, double rstep, double rplug,
                                           uint32_t zsize, double zstep) {
    if (!h || !std::isfinite(rstep) || !std::isfinite(rplug) || !std::isfinite(zstep) ||
        (uint64_t)(rsize > 1 ? rsize : 1) * (zsize > 1 ? zsize : 1) > (1ull << 31)) {
        set_error("gpe_vortex_state_set_dimensions: invalid table %u x %u", rsize, zsize);
        return GPE_EINVAL;
    }
    set_dimensions(h->vs, rsize, rstep, rplug, zsize, zstep);
    return GPE_OK;
}

// setRange (VortexState.cpp:42-47): sizes rounded to the nearest count,
// steps stretched so the table spans [rmin, rmax) x [0, zmax) exactly.
gpe_status gpe_vortex_state_set_range(gpe_vortex_state *h, double rmin, double rmax, double zmax, double rstep,
                                      double zstep) {
    if (!h || !(rstep > 0.0) || !(zstep > 0.0) || !(rmax > rmin) || !(zmax > 0.0) || !std::isfinite(rmax) ||
        !std::isfinite(zmax)) {
        set_error("gpe_vortex_state_set_range: need rmax > rmin, zmax > 0 and positive steps");
        return GPE_EINVAL;
    }
    const double nr = (rmax - rmin) / rstep + 0.5, nz = zmax / zstep + 0.5;
    if (nr < 1.0 || nz < 1.0 || nr * nz > (double)(1ull << 31)) {
        set_error("gpe_vortex_state_set_range: table of %.0f x %.0f entries", nr, nz);
        return GPE_EINVAL;
    }
    const uint32_t rsize = (uint32_t)nr, zsize = (uint32_t)nz;
    set_dimensions(h->vs, rsize, (rmax - rmin) / (double)rsize, rmin, zsize, zmax / (double)zsize);
    return GPE_OK;
}

// initialize (VortexState.cpp:98-107): mu and f sampled from src on this table.
gpe_status gpe_vortex_state_initialize(gpe_vortex_state *h, const gpe_vortex_state *src) {
    if (!h || !src) {
        set_error("gpe_vortex_state_initialize: null argument");
        return GPE_EINVAL;
    }
    VortexState &v = h->vs;
    v.mu = src->vs.mu;
    size_t ii = 0;
    for (uint32_t i = 0; i < v.rsize; i++) {
        const double r = v.rplug + i * v.rstep;
        for (uint32_t j = 0; j < v.zsize; j++) v.f[ii++] = src->vs.function(r, j * v.zstep);
    }
    return GPE_OK;
}

// iterateSolution (VortexState.cpp:109-196).  Each iteration is one sweep,
// then mu = <f,Hf>/<f,f> and f /= sqrt(<f,f>) (the r = rplug row stays 0);
// it stops after the iteration whose sweep changed f by less than eps
// (sum of squared changes times ring volumes), or after maxiters.  Finally
// f is made non-negative.  Where the reference prints progress (every 1000
// iterations) `progress` is called instead; nothing is printed.
gpe_status gpe_vortex_state_iterate(gpe_vortex_state *h, gpe_rz_potential V, double trap_ratio, void *user,
                                    uint32_t maxiters, double eps, gpe_relax_progress progress, void *progress_user,
                                    gpe_relax_report *report) {
    if (!h) {
        set_error("gpe_vortex_state_iterate: null state");
        return GPE_EINVAL;
    }
    VortexState &v = h->vs;
    const RzPotential pot{V, trap_ratio, user};
    const RzOperator op(v);
    for (uint32_t j = 0; j < v.zsize; j++) v.f[j] = 0.0;
    uint32_t it = 0;
    double diff = 0.0, prob = 0.0, prob2 = 0.0;
    bool converged = false;
    // A caller's potential runs on every worker; in automatic mode it is only
    // called from this thread.
    const int workers = (V && h->threads == 0) ? 1 : relax_workers(v, h->threads);
    if (workers > 1) {
        RelaxTeam team(v, op, pot, workers);
        std::vector<std::thread> pool;
        for (int t = 1; t < workers; t++)
            pool.emplace_back([&team, t, maxiters, eps] { team.run(t, maxiters, eps, nullptr, nullptr); });
        it = team.run(0, maxiters, eps, progress, progress_user);
        for (auto &th : pool) th.join();
        converged = team.stop.load() == 2;
        if (converged) it -= 1;  // index of the last iteration, as in the serial loop
        diff = team.diff;
        prob = team.prob;
        prob2 = team.prob2;
    } else {
        for (it = 0; it < maxiters; it++) {
            diff = relax_sweep(v, op, pot);
            rayleigh_sums(v, op, pot, prob, prob2);
            v.mu = prob2 / prob;
            const double scale = 1.0 / std::sqrt(prob);
            for (size_t ii = v.zsize; ii < v.f.size(); ii++) v.f[ii] *= scale;
            if (progress && it % 1000 == 0) progress(it, diff, prob, progress_user);
            if (diff < eps) {
                converged = true;
                break;
            }
        }
    }
    double sum = 0.0;
    for (size_t ii = v.zsize; ii < v.f.size(); ii++) sum += v.f[ii];
    if (sum < 0.0)
        for (size_t ii = v.zsize; ii < v.f.size(); ii++) v.f[ii] = -v.f[ii];
    if (report) {
        report->iterations = converged ? it + 1 : it;
        report->converged = converged ? 1u : 0u;
        report->diff = diff;
        report->prob = prob;
        report->mu = v.mu;
        report->error = residual(v, pot);
    }
    if (!std::isfinite(v.mu)) {
        set_error("gpe_vortex_state_iterate: mu became %g (relaxation diverged)", v.mu);
        return GPE_EUNSTABLE;
    }
    return GPE_OK;
}

double gpe_vortex_state_function_error(const gpe_vortex_state *h, gpe_rz_potential V, double trap_ratio,
                                       void *user) {
    return residual(h->vs, RzPotential{V, trap_ratio, user});
}

double gpe_vortex_state_function(const gpe_vortex_state *h, double r, double z) { return h->vs.function(r, z); }

// Smallest radius row holding a value above eps (VortexState.cpp:250-257).
double gpe_vortex_state_search_min_r(const gpe_vortex_state *h, double eps) {
    const VortexState &v = h->vs;
    for (size_t ii = 0; ii < v.f.size(); ii++)
        if (v.f[ii] > eps) return v.rplug + (double)(uint32_t)(ii / v.zsize) * v.rstep;
    return 1e30;
}

// integrateZ (VortexState.cpp:342-361): the Z with int_{|z|<Z} |psi|^2 = vol,
// linearly interpolated inside the z cell where the running integral
// (midpoint f between z_j and z_j+1) passes vol.
double gpe_vortex_state_integrate_z(const gpe_vortex_state *h, double vol) {
    const VortexState &v = h->vs;
    if (vol <= 0.0) return 0.0;
    const double w = 4.0 * kPi * v.rstep * v.zstep;
    double sum = 0.0;
    for (uint32_t j = 0; j < v.zsize; j++) {
        const double before = sum;
        for (uint32_t i = 1; i < v.rsize; i++) {
            const size_t ii = (size_t)i * v.zsize + j;
            const double r = v.rplug + i * v.rstep;
            const double fm = 0.5 * (v.f[ii] + (j + 1 < v.zsize ? v.f[ii + 1] : 0.0));
            sum += w * r * fm * fm;
        }
        if (sum > vol) return (j + (vol - before) / (sum - before)) * v.zstep;
    }
    return v.zsize * v.zstep;
}

// save (VortexState.cpp:74-96): the layout gpe_vortex_state_load reads.
gpe_status gpe_vortex_state_save(const gpe_vortex_state *h, const char *path) {
    if (!h || !path) {
        set_error("gpe_vortex_state_save: null argument");
        return GPE_EINVAL;
    }
    const VortexState &v = h->vs;
    std::ofstream fs(path, std::ios::binary | std::ios::trunc);
    if (!fs) {
        set_error("gpe_vortex_state_save: cannot open %s", path);
        return GPE_EIO;
    }
    fs.write((const char *)&v.kappa, 4);
    fs.write((const char *)&v.g, 8);
    fs.write((const char *)&v.rsize, 4);
    fs.write((const char *)&v.zsize, 4);
    fs.write((const char *)&v.rplug, 8);
    fs.write((const char *)&v.rstep, 8);
    fs.write((const char *)&v.zstep, 8);
    fs.write((const char *)&v.mu, 8);
    fs.write((const char *)v.f.data(), (std::streamsize)(v.f.size() * sizeof(double)));
    if (!fs) {
        set_error("gpe_vortex_state_save: write failed on %s", path);
        return GPE_EIO;
    }
    return GPE_OK;
}

// The state recipe of GPUGrossPitaevskii_BCC.cu:516-535 (LOAD_STATE_FROM_DISK
// 0): a coarse state on [0,15) x [0,35) (3-D) or [0,15) x [0,1) (2-D, one z
// row) at step 0.2 relaxed for `iters` iterations, its support measured at
// eps = 1e-5 max f, then a fine state at step 0.03 over that support,
// sampled from the coarse one and relaxed again.  extent = {minr, maxr,
// maxz} of the fine state, measured with the coarse state's eps.
gpe_status gpe_vortex_state_reference_recipe(uint32_t kappa, double g, double trap_ratio, int32_t is_3d,
                                             uint32_t iters, double eps_relax, gpe_relax_progress progress,
                                             void *progress_user, gpe_vortex_state **out, double extent[3]) {
    if (!out) {
        set_error("gpe_vortex_state_reference_recipe: null argument");
        return GPE_EINVAL;
    }
    gpe_vortex_state *coarse = nullptr, *fine = nullptr;
    gpe_status st = gpe_vortex_state_create(&coarse);
    if (st != GPE_OK) return st;
    coarse->vs.kappa = kappa;
    coarse->vs.g = g;
    st = is_3d ? gpe_vortex_state_set_range(coarse, 0.0, 15.0, 35.0, 0.2, 0.2)
               : gpe_vortex_state_set_range(coarse, 0.0, 15.0, 1.0, 0.2, 1.0);
    if (st == GPE_OK)
        st = gpe_vortex_state_iterate(coarse, nullptr, trap_ratio, nullptr, iters, eps_relax, progress, progress_user,
                                      nullptr);
    if (st != GPE_OK) {
        gpe_vortex_state_free(coarse);
        return st;
    }
    const double eps = 1e-5 * gpe_vortex_state_search_max(coarse);
    const double minr = gpe_vortex_state_search_min_r(coarse, eps);
    const double maxr = gpe_vortex_state_search_max_r(coarse, eps);
    const double maxz = gpe_vortex_state_search_max_z(coarse, eps);
    st = gpe_vortex_state_create(&fine);
    if (st == GPE_OK) {
        fine->vs.kappa = kappa;
        fine->vs.g = g;
        st = is_3d ? gpe_vortex_state_set_range(fine, minr, maxr, maxz, 0.03, 0.03)
                   : gpe_vortex_state_set_range(fine, minr, maxr, 1.0, 0.03, 1.0);
    }
    if (st == GPE_OK) st = gpe_vortex_state_initialize(fine, coarse);
    if (st == GPE_OK)
        st = gpe_vortex_state_iterate(fine, nullptr, trap_ratio, nullptr, iters, eps_relax, progress, progress_user,
                                      nullptr);
    gpe_vortex_state_free(coarse);
    if (st != GPE_OK) {
        gpe_vortex_state_free(fine);
        return st;
    }
    if (extent) {
        extent[0] = minr;
        extent[1] = gpe_vortex_state_search_max_r(fine, eps);
        extent[2] = gpe_vortex_state_search_max_z(fine, eps);
    }
    *out = fine;
    return GPE_OK;
}

}  // extern "C"
