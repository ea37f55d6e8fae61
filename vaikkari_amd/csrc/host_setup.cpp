// host_setup.cpp -- host side of the stepper: BCC dual-mesh geometry, the
// reference's VortexState initial state, the setup of integrateInTime and
// the initial-state recipes of the benchmark workloads.  Everything here is
// FP64 and runs once per solve; it is compiled with -ffp-contract=off so the
// arithmetic is the sequence written (as the reference's host code, built
// by gcc for x86-64, evaluates it).
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <fstream>
#include <new>
#include <random>
#include <algorithm>
#include <array>

#include "gpe_internal.hpp"
#include "vortex_state.hpp"
#include "../../include/vaikkari_gpe_recipes.h"

namespace gpe {

static thread_local char g_last_error[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

const int kClassPos[GPE_CLASSES][3] = {{1, 3, 7}, {7, 1, 3}, {7, 3, 1}, {7, 3, 5}, {7, 5, 3}, {1, 7, 3},
                                       {3, 7, 1}, {3, 7, 5}, {5, 7, 3}, {3, 1, 7}, {3, 5, 7}, {5, 3, 7}};

double node_measure(double b) { return std::sqrt(32.0 / 9.0) * b * b * b; }

double trap_vmax(const uint32_t dims[3], const double origin[3], double b, double ratio) {
    double vmax = 0.0;
    for (int ci = 0; ci < 2; ci++)
        for (int cj = 0; cj < 2; cj++)
            for (int ck = 0; ck < 2; ck++)
                for (int l = 0; l < GPE_CLASSES; l++) {
                    double p[3];
                    node_position(origin, b, ci ? dims[0] - 1 : 0, cj ? dims[1] - 1 : 0, ck ? dims[2] - 1 : 0, l, p);
                    vmax = std::max(vmax, trap_potential(p, ratio));
                }
    return vmax;
}

bool linear_step_stable(double dt, double b, double vmax) { return dt * (6.0 / (b * b) + vmax) < 2.0; }

void node_position(const double origin[3], double b, uint32_t i, uint32_t j, uint32_t k, int l, double p[3]) {
    const double bw = std::sqrt(8.0), sq18 = 1.0 / std::sqrt(8.0);
    const uint32_t ijk[3] = {i, j, k};
    for (int a = 0; a < 3; a++) p[a] = origin[a] + b * (ijk[a] * bw + sq18 * kClassPos[l][a]);
}

}  // namespace gpe

using namespace gpe;

extern "C" {

const char *gpe_last_error(void) { return gpe::g_last_error; }

const char *gpe_status_string(gpe_status st) {
    switch (st) {
        case GPE_OK: return "ok";
        case GPE_EINVAL: return "invalid argument";
        case GPE_ENOMEM: return "out of memory";
        case GPE_EDEVICE: return "device error";
        case GPE_EUNSTABLE: return "unstable or non-finite state";
        case GPE_EBUSY: return "budget exhausted";
        case GPE_ENOTSUP: return "not supported";
        case GPE_EIO: return "i/o error";
        case GPE_ECANCELED: return "cancelled";
        case GPE_ETIMEDOUT: return "timed out";
    }
    return "unknown status";
}

void gpe_bcc_positions(int32_t out[36]) {
    for (int l = 0; l < GPE_CLASSES; l++)
        for (int a = 0; a < 3; a++) out[3 * l + a] = kClassPos[l][a];
}

void gpe_bcc_stencil(int32_t out[192]) {
    for (int l = 0; l < GPE_CLASSES; l++)
        for (int f = 0; f < GPE_FACES; f++) {
            const Face e = bcc_face(l, f);
            int32_t *o = out + 4 * (GPE_FACES * l + f);
            o[0] = e.dx; o[1] = e.dy; o[2] = e.dz; o[3] = e.m;
        }
}

// Two dual nodes are adjacent exactly when their distance is b, i.e. when
// |q_m + 8*offset - q_l|^2 = 8 in the integer units of q/sqrt(8)*b.
int gpe_bcc_stencil_derived(int32_t out[192]) {
    int total = 0;
    for (int l = 0; l < GPE_CLASSES; l++) {
        std::vector<std::array<int, 4>> nb;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++)
                    for (int m = 0; m < GPE_CLASSES; m++) {
                        const int d0 = kClassPos[m][0] + 8 * dx - kClassPos[l][0];
                        const int d1 = kClassPos[m][1] + 8 * dy - kClassPos[l][1];
                        const int d2 = kClassPos[m][2] + 8 * dz - kClassPos[l][2];
                        if (d0 * d0 + d1 * d1 + d2 * d2 == 8) nb.push_back({m, dz, dy, dx});
                    }
        std::sort(nb.begin(), nb.end());
        for (size_t f = 0; f < nb.size() && f < (size_t)GPE_FACES; f++) {
            int32_t *o = out + 4 * (GPE_FACES * l + (int)f);
            o[0] = nb[f][3]; o[1] = nb[f][2]; o[2] = nb[f][1]; o[3] = nb[f][0];
        }
        total += (int)nb.size();
    }
    return total;
}

// VortexState::load, VortexState.cpp:49-72: packed little-endian fields
// (uint kappa, double g, uint rsize, uint zsize, double rplug, rstep, zstep,
// mu) followed by rsize*zsize doubles of f.
gpe_status gpe_vortex_state_load(const char *path, gpe_vortex_state **out) {
    if (!path || !out) {
        set_error("gpe_vortex_state_load: null argument");
        return GPE_EINVAL;
    }
    std::ifstream fs(path, std::ios::binary);
    if (!fs) {
        set_error("gpe_vortex_state_load: cannot open %s", path);
        return GPE_EIO;
    }
    auto *h = new gpe_vortex_state();
    VortexState &v = h->vs;
    fs.read((char *)&v.kappa, 4);
    fs.read((char *)&v.g, 8);
    fs.read((char *)&v.rsize, 4);
    fs.read((char *)&v.zsize, 4);
    fs.read((char *)&v.rplug, 8);
    fs.read((char *)&v.rstep, 8);
    fs.read((char *)&v.zstep, 8);
    fs.read((char *)&v.mu, 8);
    if (!fs || v.rsize == 0 || v.zsize == 0 || v.rsize > (1u << 20) || v.zsize > (1u << 20)) {
        delete h;
        set_error("gpe_vortex_state_load: bad header in %s", path);
        return GPE_EIO;
    }
    v.f.resize((size_t)v.rsize * v.zsize);
    fs.read((char *)v.f.data(), (std::streamsize)(v.f.size() * sizeof(double)));
    if (!fs) {
        delete h;
        set_error("gpe_vortex_state_load: truncated table in %s", path);
        return GPE_EIO;
    }
    *out = h;
    return GPE_OK;
}

void gpe_vortex_state_free(gpe_vortex_state *vs) { delete vs; }

gpe_status gpe_vortex_state_from_table(uint32_t kappa, double g, uint32_t rsize, uint32_t zsize, double rplug,
                                       double rstep, double zstep, double mu, const double *f,
                                       gpe_vortex_state **out) {
    if (!f || !out || rsize == 0 || zsize == 0 || rsize > (1u << 20) || zsize > (1u << 20) || !(rstep > 0) ||
        !(zstep > 0) || !std::isfinite(rplug) || !std::isfinite(g) || !std::isfinite(mu)) {
        set_error("gpe_vortex_state_from_table: invalid argument");
        return GPE_EINVAL;
    }
    auto *h = new (std::nothrow) gpe_vortex_state();
    if (!h) return GPE_ENOMEM;
    VortexState &v = h->vs;
    v.kappa = kappa;
    v.g = g;
    v.rsize = rsize;
    v.zsize = zsize;
    v.rplug = rplug;
    v.rstep = rstep;
    v.zstep = zstep;
    v.mu = mu;
    v.f.assign(f, f + (size_t)rsize * zsize);
    *out = h;
    return GPE_OK;
}

gpe_status gpe_vortex_state_table(const gpe_vortex_state *h, double *f) {
    if (!h || !f) {
        set_error("gpe_vortex_state_table: null argument");
        return GPE_EINVAL;
    }
    std::copy(h->vs.f.begin(), h->vs.f.end(), f);
    return GPE_OK;
}

void gpe_vortex_state_params(const gpe_vortex_state *h, double out[8]) {
    const VortexState &v = h->vs;
    out[0] = v.kappa; out[1] = v.g; out[2] = v.rsize; out[3] = v.zsize;
    out[4] = v.rplug; out[5] = v.rstep; out[6] = v.zstep; out[7] = v.mu;
}

void gpe_vortex_state_psi(const gpe_vortex_state *h, double x, double y, double z, double out[2]) {
    h->vs.psi(x, y, z, out);
}

// getPhase(t) = exp(-i mu t) (VortexState.cpp:198-202).
void gpe_vortex_state_phase(const gpe_vortex_state *h, double time, double out[2]) {
    const double ph = time * h->vs.mu;
    out[0] = std::cos(ph);
    out[1] = -std::sin(ph);
}

double gpe_vortex_state_search_max(const gpe_vortex_state *h) {
    double m = 0.0;
    for (double x : h->vs.f) m = x > m ? x : m;
    return m;
}

// Largest radius row holding a value above eps (VortexState.cpp:259-266).
double gpe_vortex_state_search_max_r(const gpe_vortex_state *h, double eps) {
    const VortexState &v = h->vs;
    for (size_t ii = v.f.size(); ii-- > 0;)
        if (v.f[ii] > eps) return v.rplug + (double)(uint32_t)(ii / v.zsize) * v.rstep;
    return 0.0;
}

// Largest z column holding a value above eps (VortexState.cpp:268-280).
double gpe_vortex_state_search_max_z(const gpe_vortex_state *h, double eps) {
    const VortexState &v = h->vs;
    for (uint32_t j = v.zsize; j-- > 0;)
        for (uint32_t i = 0; i < v.rsize; i++)
            if (v.f[(size_t)i * v.zsize + j] > eps) return (double)j * v.zstep;
    return 0.0;
}

// Mean relative curvature <-lap psi, psi> on the (r,z) grid with the
// reference's finite differences and weights (VortexState.cpp:313-340).
double gpe_vortex_state_curvature(const gpe_vortex_state *h) {
    const VortexState &v = h->vs;
    const double two_pi = 6.2831853071795864770;
    const double w0 = two_pi * v.rstep * v.zstep;
    const double kk = (double)(v.kappa * v.kappa);
    const double cr = 1.0 / (v.rstep * v.rstep);
    const double cz = (v.zsize > 1 ? 1.0 / (v.zstep * v.zstep) : 0.0);
    double curv = 0.0;
    for (uint32_t i = 1; i < v.rsize; i++) {
        const double r = v.rplug + i * v.rstep;
        const double diag = kk / (r * r) + 2.0 * (cr + cz);
        const double grad = 0.5 / (r * v.rstep);
        double dv = w0 * r;
        const double *col = v.f.data() + (size_t)i * v.zsize;
        for (uint32_t j = 0; j < v.zsize; j++) {
            const double up_r = (i + 1 < v.rsize ? col[j + v.zsize] : 0.0);
            const double dn_r = col[(ptrdiff_t)j - (ptrdiff_t)v.zsize];
            const double up_z = (j + 1 < v.zsize ? col[j + 1] : 0.0);
            const double dn_z = (j > 0 ? col[j - 1] : up_z);  // mirror at z = 0
            const double c = diag * col[j] + grad * (dn_r - up_r) - cr * (dn_r + up_r) - cz * (dn_z + up_z);
            curv += c * col[j] * dv;
            if (j == 0) dv *= 2.0;  // the z = 0 row stands for one half-plane only
        }
    }
    return curv;
}

void gpe_reference_dims(double b, const double minp[3], const double maxp[3], uint32_t dims[3], double origin[3]) {
    const double bw = std::sqrt(8.0);
    for (int a = 0; a < 3; a++) dims[a] = (uint32_t)((maxp[a] - minp[a]) / (b * bw)) + 1;
    for (int a = 0; a < 3; a++) origin[a] = ((minp[a] + maxp[a]) - b * (bw * dims[a])) * 0.5;
}

gpe_status gpe_reference_setup(const gpe_vortex_state *h, double b, const double minp[3], const double maxp[3],
                               double period, double ratio, uint32_t seed, gpe_reference_run *run, double *even,
                               double *odd, double *V) {
    if (!h || !run || !(b > 0) || !(period > 0)) {
        set_error("gpe_reference_setup: invalid argument");
        return GPE_EINVAL;
    }
    const VortexState &v = h->vs;
    gpe_reference_dims(b, minp, maxp, run->dims, run->origin);
    const uint32_t X = run->dims[0], Y = run->dims[1], Z = run->dims[2];
    const size_t n = (size_t)X * Y * Z * GPE_CLASSES;
    // psi0 and V at every node; maxpot = max(V + g|psi0|^2)
    // (GPUGrossPitaevskii_BCC.cu:249-266).
    std::vector<double> psi0(2 * n), pot(n);
    double maxpot = 0.0;
    size_t ii = 0;
    for (uint32_t k = 0; k < Z; k++)
        for (uint32_t j = 0; j < Y; j++)
            for (uint32_t i = 0; i < X; i++)
                for (int l = 0; l < GPE_CLASSES; l++, ii++) {
                    double p[3];
                    node_position(run->origin, b, i, j, k, l, p);
                    v.psi(p[0], p[1], p[2], &psi0[2 * ii]);
                    pot[ii] = trap_potential(p, ratio);
                    const double e = pot[ii] + v.g * (psi0[2 * ii] * psi0[2 * ii] + psi0[2 * ii + 1] * psi0[2 * ii + 1]);
                    if (e > maxpot) maxpot = e;
                }
    // Time step: steps = floor(T*(maxpot + lapfac0)) + 1, dt = T/steps
    // (GPUGrossPitaevskii_BCC.cu:291-299).
    const double lapfac = -0.5 * kHodge / (b * b);
    const double lapfac0 = (double)GPE_FACES * (-lapfac);
    const uint32_t steps = (uint32_t)(period * (maxpot + lapfac0)) + 1;
    const double dt = period / (double)steps;
    run->steps_per_iteration = steps;
    run->dt = dt;
    run->g = v.g;
    run->maxpot = maxpot;
    run->lapfac_dt = lapfac * dt;
    run->lapfac0_dt = lapfac0 * dt;
    if (V) std::memcpy(V, pot.data(), n * sizeof(double));
    if (even || odd) {
        // even = psi0 * (1 + 0.01 c), c uniform in the unit disk from MT19937
        // seeded with `seed`, drawn in node order; odd = even * exp(+i mu dt/2)
        // (GPUGrossPitaevskii_BCC.cu:325-345, Random.cpp:77-86).  std::mt19937
        // is the same generator as the reference's Random (init_genrand,
        // twist and tempering of MT19937).
        std::mt19937 rng(seed);
        const double ph = (-0.5 * dt) * v.mu;
        const double cr = std::cos(ph), ci = -std::sin(ph);
        const double scale31 = 1.0 / 2147483648.0;
        for (ii = 0; ii < n; ii++) {
            double cx, cy;
            do {
                cx = scale31 * (double)(int32_t)rng();
                cy = scale31 * (double)(int32_t)rng();
            } while (cx * cx + cy * cy > 1.0);
            const double nr = 0.01 * cx + 1.0, ni = 0.01 * cy;
            const double pr = psi0[2 * ii], pi = psi0[2 * ii + 1];
            const double er = pr * nr - pi * ni, ei = pi * nr + pr * ni;
            if (even) {
                even[2 * ii] = er;
                even[2 * ii + 1] = ei;
            }
            if (odd) {
                odd[2 * ii] = cr * er - ci * ei;
                odd[2 * ii + 1] = ci * er + cr * ei;
            }
        }
    }
    return GPE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Initial-state recipes of the benchmark workloads (this library's own
// definitions; the reference has none of them).
// ---------------------------------------------------------------------

namespace gpe {

// mt19937_64 -> N(0,1) pairs by explicit Box-Muller (std::normal_distribution
// is implementation-defined, so it is not used).
struct NormalPairs {
    std::mt19937_64 rng;
    explicit NormalPairs(uint64_t seed) : rng(seed) {}
    void next(double &a, double &c) {
        const double two_pi = 6.2831853071795864770;
        const double u1 = ((double)(rng() >> 11) + 1.0) * (1.0 / 9007199254740992.0);  // (0,1]
        const double u2 = (double)(rng() >> 11) * (1.0 / 9007199254740992.0);          // [0,1)
        const double r = std::sqrt(-2.0 * std::log(u1));
        a = r * std::cos(two_pi * u2);
        c = r * std::sin(two_pi * u2);
    }
};

static size_t grid_nodes(const uint32_t dims[3]) { return (size_t)dims[0] * dims[1] * dims[2] * GPE_CLASSES; }

template <class F>
static void for_nodes(const uint32_t dims[3], const double origin[3], double b, F &&fn) {
    size_t n = 0;
    for (uint32_t k = 0; k < dims[2]; k++)
        for (uint32_t j = 0; j < dims[1]; j++)
            for (uint32_t i = 0; i < dims[0]; i++)
                for (int l = 0; l < GPE_CLASSES; l++, n++) {
                    double p[3];
                    node_position(origin, b, i, j, k, l, p);
                    fn(n, p);
                }
}

static double field_norm(const double *psi, size_t n, double b) {
    // Neumaier-compensated sum of |psi|^2 times the node measure.
    double s = 0.0, c = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double x = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        const double t = s + x;
        c += std::fabs(s) >= std::fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    return node_measure(b) * (s + c);
}

}  // namespace gpe

extern "C" {

void gpe_centered_origin(const uint32_t dims[3], double b, double origin[3]) {
    const double a = std::sqrt(8.0) * b;
    for (int d = 0; d < 3; d++) origin[d] = -0.5 * a * dims[d];
}

void gpe_node_positions(const uint32_t dims[3], const double origin[3], double b, double *out) {
    for_nodes(dims, origin, b, [&](size_t n, const double *p) {
        out[3 * n] = p[0];
        out[3 * n + 1] = p[1];
        out[3 * n + 2] = p[2];
    });
}

void gpe_trap_potential(const uint32_t dims[3], const double origin[3], double b, double ratio, double *V) {
    for_nodes(dims, origin, b, [&](size_t n, const double *p) { V[n] = trap_potential(p, ratio); });
}

double gpe_field_norm(const uint32_t dims[3], double b, const double *psi) {
    return field_norm(psi, grid_nodes(dims), b);
}

gpe_status gpe_recipe_gaussian(const uint32_t dims[3], const double origin[3], double b, double sigma,
                               double noise, uint64_t seed, double *psi) {
    if (!dims || !origin || !psi || !(b > 0) || !(sigma > 0)) {
        set_error("gpe_recipe_gaussian: invalid argument");
        return GPE_EINVAL;
    }
    const size_t n = grid_nodes(dims);
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    for_nodes(dims, origin, b, [&](size_t i, const double *p) {
        psi[2 * i] = std::exp(-(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) * inv2s2);
        psi[2 * i + 1] = 0.0;
    });
    const double nrm = field_norm(psi, n, b);
    if (!(nrm > 0)) {
        set_error("gpe_recipe_gaussian: zero norm (grid does not cover the Gaussian)");
        return GPE_EINVAL;
    }
    const double sc = 1.0 / std::sqrt(nrm);
    NormalPairs rnd(seed);
    for (size_t i = 0; i < n; i++) {
        double xr, xi;
        rnd.next(xr, xi);
        const double a = psi[2 * i] * sc;
        psi[2 * i] = a * (1.0 + noise * xr);
        psi[2 * i + 1] = a * (noise * xi);
    }
    return GPE_OK;
}

gpe_status gpe_recipe_thomas_fermi(const uint32_t dims[3], const double origin[3], double b, double g,
                                   double ratio, double *psi, double *mu_out) {
    if (!dims || !origin || !psi || !(b > 0) || !(g > 0)) {
        set_error("gpe_recipe_thomas_fermi: invalid argument");
        return GPE_EINVAL;
    }
    const size_t n = grid_nodes(dims);
    std::vector<double> V(n);
    gpe_trap_potential(dims, origin, b, ratio, V.data());
    const double w = node_measure(b);
    auto norm_at = [&](double mu) {
        double s = 0.0;
        for (size_t i = 0; i < n; i++) s += std::max(mu - V[i], 0.0);
        return w * s / g;
    };
    double lo = 0.0, hi = 1.0;
    while (norm_at(hi) < 1.0 && hi < 1e12) hi *= 2.0;
    for (int it = 0; it < 200; it++) {
        const double mid = 0.5 * (lo + hi);
        (norm_at(mid) < 1.0 ? lo : hi) = mid;
    }
    const double mu = 0.5 * (lo + hi);
    for (size_t i = 0; i < n; i++) {
        psi[2 * i] = std::sqrt(std::max(mu - V[i], 0.0) / g);
        psi[2 * i + 1] = 0.0;
    }
    const double sc = 1.0 / std::sqrt(field_norm(psi, n, b));
    for (size_t i = 0; i < 2 * n; i++) psi[i] *= sc;
    if (mu_out) *mu_out = mu;
    return GPE_OK;
}

gpe_status gpe_recipe_vortex_line(const uint32_t dims[3], const double origin[3], double b, double x0, double y0,
                                  double *psi) {
    if (!dims || !origin || !psi) {
        set_error("gpe_recipe_vortex_line: invalid argument");
        return GPE_EINVAL;
    }
    for_nodes(dims, origin, b, [&](size_t i, const double *p) {
        const double dx = p[0] - x0, dy = p[1] - y0;
        const double rho = std::sqrt(dx * dx + dy * dy);
        const double amp = std::sqrt(psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1]);
        if (rho < 1e-30) {
            psi[2 * i] = psi[2 * i + 1] = 0.0;
        } else {
            psi[2 * i] = amp * dx / rho;
            psi[2 * i + 1] = amp * dy / rho;
        }
    });
    return GPE_OK;
}

gpe_status gpe_recipe_vortex_ring(const uint32_t dims[3], const double origin[3], double b, const double center[3],
                                  const double axis[3], double radius, double *psi) {
    if (!dims || !origin || !psi || !center || !axis || !(radius > 0)) {
        set_error("gpe_recipe_vortex_ring: invalid argument");
        return GPE_EINVAL;
    }
    const double an = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (!(an > 0)) {
        set_error("gpe_recipe_vortex_ring: zero axis");
        return GPE_EINVAL;
    }
    const double ax[3] = {axis[0] / an, axis[1] / an, axis[2] / an};
    for_nodes(dims, origin, b, [&](size_t i, const double *p) {
        const double d[3] = {p[0] - center[0], p[1] - center[1], p[2] - center[2]};
        const double zl = d[0] * ax[0] + d[1] * ax[1] + d[2] * ax[2];
        const double r[3] = {d[0] - zl * ax[0], d[1] - zl * ax[1], d[2] - zl * ax[2]};
        const double rho = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) - radius;
        const double den = std::sqrt(rho * rho + zl * zl);
        const double amp = std::sqrt(psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1]);
        if (den < 1e-30) {
            psi[2 * i] = psi[2 * i + 1] = 0.0;
        } else {
            psi[2 * i] = amp * rho / den;
            psi[2 * i + 1] = amp * zl / den;
        }
    });
    return GPE_OK;
}

gpe_status gpe_recipe_phase_shift(const uint32_t dims[3], const double *even, double mu, double dt, double *odd) {
    // odd = even * exp(+i mu dt/2), the start of the leapfrog
    // (GPUGrossPitaevskii_BCC.cu:325,342-343).
    if (!dims || !even || !odd) {
        set_error("gpe_recipe_phase_shift: invalid argument");
        return GPE_EINVAL;
    }
    const double ph = (-0.5 * dt) * mu;
    const double cr = std::cos(ph), ci = -std::sin(ph);
    const size_t n = grid_nodes(dims);
    for (size_t i = 0; i < n; i++) {
        const double er = even[2 * i], ei = even[2 * i + 1];
        odd[2 * i] = cr * er - ci * ei;
        odd[2 * i + 1] = ci * er + cr * ei;
    }
    return GPE_OK;
}

}  // extern "C"
