// ref_shim.cpp -- C entry points over the reference's own host C++ (TEST
// INFRASTRUCTURE ONLY).  oracle/Makefile compiles this file together with
// koodit/GrossPitaevskii/VortexState.cpp and jukan_koodit/Types/Random.cpp
// straight from the reference tree into oracle/_ref/libvsref.so.  Tests use
// it to pin the C restatement in oracle/gpe_oracle.c and the product's host
// setup; nothing else loads it.
#include "VortexState.hpp"
#include "../../jukan_koodit/Types/Random.hpp"
#include <stdint.h>

extern "C" {

void *ref_vs_load(const char *path) {
    VortexState *s = new VortexState();
    if (!s->load(path)) { delete s; return nullptr; }
    return s;
}
void ref_vs_free(void *h) { delete static_cast<VortexState *>(h); }

void ref_vs_psi(void *h, double x, double y, double z, double *out) {
    const Complex c = static_cast<VortexState *>(h)->getPsi(Vector3(x, y, z));
    out[0] = c.r;
    out[1] = c.i;
}
double ref_vs_function(void *h, double r, double z) { return static_cast<VortexState *>(h)->getFunction(r, z); }
void ref_vs_phase(void *h, double t, double *out) {
    const Complex c = static_cast<VortexState *>(h)->getPhase(t);
    out[0] = c.r;
    out[1] = c.i;
}
double ref_vs_search_max(void *h) { return static_cast<VortexState *>(h)->searchFunctionMax(); }
double ref_vs_search_max_r(void *h, double eps) { return static_cast<VortexState *>(h)->searchMaxR(eps); }
double ref_vs_search_max_z(void *h, double eps) { return static_cast<VortexState *>(h)->searchMaxZ(eps); }
double ref_vs_curvature(void *h) { return static_cast<VortexState *>(h)->integrateCurvature(); }
double ref_vs_mu(void *h) { return static_cast<VortexState *>(h)->getMu(); }
double ref_vs_g(void *h) { return static_cast<VortexState *>(h)->getG(); }
double ref_vs_kappa(void *h) { return static_cast<VortexState *>(h)->getKappa(); }

void ref_random_uints(uint32_t seed, long long n, uint32_t *out) {
    Random r(seed);
    for (long long i = 0; i < n; i++) out[i] = r.getUint();
}
void ref_random_circles(uint32_t seed, long long n, double *out) {
    Random r(seed);
    for (long long i = 0; i < n; i++) {
        const Vector2 v = r.getUniformCircle();
        out[2 * i] = v.x;
        out[2 * i + 1] = v.y;
    }
}

}  // extern "C"
