/*
 * primal_oracle.c -- CPU restatement of the 14-point primal BCC stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  Part of liboracle.so (oracle/Makefile): only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker.
 * The product library never links or calls anything in oracle/.
 *
 * No reference stepper uses this operator.  It is the DEC Laplacian of
 * 0-forms on the BCC lattice that the reference's mesh library builds
 * (DelaunayMesh::createBccGrid, jukan_koodit/Mesh/DelaunayMesh.cpp:178,
 * with Mesh::getEdgeHodge / getNodeHodge, Mesh/Mesh.cpp:1479-1489): for
 * cube edge a, sites at the cube corners and centres, 8 diagonal edges of
 * length sqrt(3)a/2 with Hodge weight 3a/8, 6 axial edges of length a with
 * weight a/8, dual volume a^3/2, so
 *   (L psi)_i = (1/a^2) [ 3/4 sum_8 (psi_j - psi_i) + 1/4 sum_6 (psi_j - psi_i) ].
 * tests/test_primal_pins.py checks those weights against the reference's
 * own mesh code (oracle/_ref) and L against closed forms.
 *
 * Hamiltonian and integrator as in the reference stepper
 * (GPUGrossPitaevskii_BCC.cu:205-215, 475-481), with the neighbour sums in
 * the fixed order used by the GPU kernel:
 *   s = cD*(((((((d0+d1)+d2)+d3)+d4)+d5)+d6)+d7) + cA*(((((a0+a1)+a2)+a3)+a4)+a5)
 *       + ((c0 + V) + g|psi|^2) psi,           all coefficients times dt,
 *   real time: next += (s.y, -s.x);  imaginary time: out = psi - s,
 * cD = -3/(8a^2), cA = -1/(8a^2), c0 = 15/(4a^2).
 * Diagonal order: dz, dy, dx in {-1,0} for a corner (offsets of the centre
 * sites), {0,+1} for a centre; axial order -x, +x, -y, +y, -z, +z.
 *
 * Storage: [k][j][i][s] complex, s = 0 corner (position p0 + a(i,j,k)),
 * s = 1 centre (p0 + a(i+1/2, j+1/2, k+1/2)); zero outside the box.
 */
#include <math.h>
#include <stddef.h>

/* Built with -ffp-contract=off: every expression is evaluated as written. */

static long long pidx(const int dims[3], int i, int j, int k, int s) {
    return ((((long long)k * dims[1] + j) * dims[0] + i) * 2 + s);
}

static int inside(const int dims[3], int i, int j, int k) {
    return i >= 0 && i < dims[0] && j >= 0 && j < dims[1] && k >= 0 && k < dims[2];
}

/* value at (i, j, k, s) with the zero halo */
static void get(const double *f, const int dims[3], int i, int j, int k, int s, double *re, double *im) {
    if (!inside(dims, i, j, k)) {
        *re = 0.0;
        *im = 0.0;
        return;
    }
    const long long e = pidx(dims, i, j, k, s);
    *re = f[2 * e];
    *im = f[2 * e + 1];
}

/* coef = {cD*dt, cA*dt, c0*dt, g*dt}; pot_dt: V*dt per site ([k][j][i][s]).
 * imag = 0: next += -i s (in place); imag = 1: next = prev*scale - s with
 * s from prev*scale.  Returns sum |next|^2 (imaginary time). */
double or_primal_half_update(double *next, const double *prev, const double *pot_dt, const int dims[3],
                             const double coef[4], int imag, double scale) {
    double norm = 0.0;
    for (int k = 0; k < dims[2]; k++)
        for (int j = 0; j < dims[1]; j++)
            for (int i = 0; i < dims[0]; i++)
                for (int s = 0; s < 2; s++) {
                    const int o = s == 0 ? -1 : 0; /* diagonal offsets {o, o+1} */
                    double nr[8], ni[8];
                    int q = 0;
                    for (int dz = 0; dz < 2; dz++)
                        for (int dy = 0; dy < 2; dy++)
                            for (int dx = 0; dx < 2; dx++, q++) {
                                get(prev, dims, i + o + dx, j + o + dy, k + o + dz, 1 - s, &nr[q], &ni[q]);
                                if (imag) {
                                    nr[q] = nr[q] * scale;
                                    ni[q] = ni[q] * scale;
                                }
                            }
                    double dr = nr[0], di = ni[0];
                    for (q = 1; q < 8; q++) {
                        dr = dr + nr[q];
                        di = di + ni[q];
                    }
                    const int ax[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
                    for (q = 0; q < 6; q++) {
                        get(prev, dims, i + ax[q][0], j + ax[q][1], k + ax[q][2], s, &nr[q], &ni[q]);
                        if (imag) {
                            nr[q] = nr[q] * scale;
                            ni[q] = ni[q] * scale;
                        }
                    }
                    double ar = nr[0], ai = ni[0];
                    for (q = 1; q < 6; q++) {
                        ar = ar + nr[q];
                        ai = ai + ni[q];
                    }
                    const long long e = pidx(dims, i, j, k, s);
                    double pr = prev[2 * e], pi = prev[2 * e + 1];
                    if (imag) {
                        pr = pr * scale;
                        pi = pi * scale;
                    }
                    const double nsq = pr * pr + pi * pi;
                    const double diagc = (coef[2] + pot_dt[e]) + coef[3] * nsq;
                    const double sr = (coef[0] * dr + coef[1] * ar) + diagc * pr;
                    const double si = (coef[0] * di + coef[1] * ai) + diagc * pi;
                    if (imag) {
                        const double orr = pr - sr, oi = pi - si;
                        next[2 * e] = orr;
                        next[2 * e + 1] = oi;
                        norm += orr * orr + oi * oi;
                    } else {
                        next[2 * e] = next[2 * e] + si;
                        next[2 * e + 1] = next[2 * e + 1] + (-sr);
                    }
                }
    return norm;
}

/* nsteps leapfrog steps: odd from even, then even from odd. */
void or_primal_leapfrog(double *even, double *odd, const double *pot_dt, const int dims[3], const double coef[4],
                        long long nsteps) {
    for (long long t = 0; t < nsteps; t++) {
        or_primal_half_update(odd, even, pot_dt, dims, coef, 0, 1.0);
        or_primal_half_update(even, odd, pot_dt, dims, coef, 0, 1.0);
    }
}

/* Imaginary time: psi <- psi - dtau H psi, then psi /= sqrt(N), N = w sum|psi|^2,
 * w = a^3/2; work: scratch of the same size.  Returns the last N. */
double or_primal_imag_steps(double *psi, double *work, const double *pot_dt, const int dims[3], const double coef[4],
                            double w, long long nsteps) {
    const long long n = 2LL * dims[0] * dims[1] * dims[2];
    double N = 0.0;
    for (long long t = 0; t < nsteps; t++) {
        const double s = or_primal_half_update(work, psi, pot_dt, dims, coef, 1, 1.0);
        N = w * s;
        const double sc = 1.0 / sqrt(N);
        for (long long e = 0; e < 2 * n; e++) psi[e] = work[e] * sc;
    }
    return N;
}

/* out = {N, K, P, I}: N = w sum|psi|^2, K = w sum 1/2 Re(conj(psi)(-L psi)),
 * P = w sum V|psi|^2, I = w sum g|psi|^4, with lD = 3/(4a^2), lA = 1/(4a^2)
 * the weights of L and V unscaled.  Plain double sums in site order. */
void or_primal_observables(const double *psi, const double *V, const int dims[3], double a, double g, double out[4]) {
    const double lD = 0.75 / (a * a), lA = 0.25 / (a * a), w = 0.5 * a * a * a;
    double N = 0.0, K = 0.0, P = 0.0, I = 0.0;
    for (int k = 0; k < dims[2]; k++)
        for (int j = 0; j < dims[1]; j++)
            for (int i = 0; i < dims[0]; i++)
                for (int s = 0; s < 2; s++) {
                    const long long e = pidx(dims, i, j, k, s);
                    const double pr = psi[2 * e], pi = psi[2 * e + 1];
                    const int o = s == 0 ? -1 : 0;
                    double lr = 0.0, li = 0.0;
                    for (int dz = 0; dz < 2; dz++)
                        for (int dy = 0; dy < 2; dy++)
                            for (int dx = 0; dx < 2; dx++) {
                                double r, m;
                                get(psi, dims, i + o + dx, j + o + dy, k + o + dz, 1 - s, &r, &m);
                                lr += lD * (r - pr);
                                li += lD * (m - pi);
                            }
                    const int ax[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
                    for (int q = 0; q < 6; q++) {
                        double r, m;
                        get(psi, dims, i + ax[q][0], j + ax[q][1], k + ax[q][2], s, &r, &m);
                        lr += lA * (r - pr);
                        li += lA * (m - pi);
                    }
                    const double nsq = pr * pr + pi * pi;
                    N += nsq;
                    K += 0.5 * -(pr * lr + pi * li);
                    P += V[e] * nsq;
                    I += g * nsq * nsq;
                }
    out[0] = w * N;
    out[1] = w * K;
    out[2] = w * P;
    out[3] = w * I;
}

/* (L f)_i for a real field f ([k][j][i][s]) on all sites, zero halo. */
void or_primal_laplacian(const double *f, double *out, const int dims[3], double a) {
    const double lD = 0.75 / (a * a), lA = 0.25 / (a * a);
    for (int k = 0; k < dims[2]; k++)
        for (int j = 0; j < dims[1]; j++)
            for (int i = 0; i < dims[0]; i++)
                for (int s = 0; s < 2; s++) {
                    const long long e = pidx(dims, i, j, k, s);
                    const int o = s == 0 ? -1 : 0;
                    double acc = 0.0;
                    for (int dz = 0; dz < 2; dz++)
                        for (int dy = 0; dy < 2; dy++)
                            for (int dx = 0; dx < 2; dx++) {
                                const int ii = i + o + dx, jj = j + o + dy, kk = k + o + dz;
                                const double v = inside(dims, ii, jj, kk) ? f[pidx(dims, ii, jj, kk, 1 - s)] : 0.0;
                                acc += lD * (v - f[e]);
                            }
                    const int ax[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
                    for (int q = 0; q < 6; q++) {
                        const int ii = i + ax[q][0], jj = j + ax[q][1], kk = k + ax[q][2];
                        const double v = inside(dims, ii, jj, kk) ? f[pidx(dims, ii, jj, kk, s)] : 0.0;
                        acc += lA * (v - f[e]);
                    }
                    out[e] = acc;
                }
}
