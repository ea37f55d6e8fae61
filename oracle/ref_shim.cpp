// ref_shim.cpp -- C entry points over the reference's own host C++ (TEST
// INFRASTRUCTURE ONLY).  oracle/Makefile compiles this file together with
// koodit/GrossPitaevskii/VortexState.cpp, jukan_koodit/Types/Random.cpp and
// jukan_koodit/Output/{Picture,Text}.cpp straight from the reference tree
// into oracle/_ref/libvsref.so.  Tests use
// it to pin the C restatement in oracle/gpe_oracle.c and the product's host
// setup; nothing else loads it.
#include "VortexState.hpp"
#include "../../jukan_koodit/Types/Random.hpp"
#include "../../jukan_koodit/Output/Picture.hpp"
#include "../../jukan_koodit/Output/Text.hpp"
#include <fstream>
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

// Generating a state with the reference's own code.  The relaxation takes a
// plain function pointer, so the trap ratio of 0.5(r^2 + ratio^2 z^2) (the
// form of potentialRZ, GPUGrossPitaevskii_BCC.cu:28-31) is a file-scope value.
static double g_ratio = 0.1;
static double shim_trap(const double r, const double z) { return 0.5 * (r * r + g_ratio * g_ratio * z * z); }

void *ref_vs_create(void) { return new VortexState(); }
void ref_vs_set_physics(void *h, unsigned kappa, double g) {
    static_cast<VortexState *>(h)->setKappa(kappa);
    static_cast<VortexState *>(h)->setG(g);
}
void ref_vs_set_dimensions(void *h, unsigned rsize, double rstep, double rplug, unsigned zsize, double zstep) {
    static_cast<VortexState *>(h)->setDimensions(rsize, rstep, rplug, zsize, zstep);
}
void ref_vs_set_range(void *h, double rmin, double rmax, double zmax, double rstep, double zstep) {
    static_cast<VortexState *>(h)->setRange(rmin, rmax, zmax, rstep, zstep);
}
void ref_vs_initialize(void *h, void *src) {
    static_cast<VortexState *>(h)->initialize(*static_cast<VortexState *>(src));
}
void ref_vs_iterate(void *h, double ratio, unsigned maxiters, double eps) {
    g_ratio = ratio;
    static_cast<VortexState *>(h)->iterateSolution(shim_trap, maxiters, eps);
}
double ref_vs_error(void *h, double ratio) {
    g_ratio = ratio;
    return static_cast<VortexState *>(h)->integrateFunctionError(shim_trap);
}
double ref_vs_search_min_r(void *h, double eps) { return static_cast<VortexState *>(h)->searchMinR(eps); }
double ref_vs_integrate_z(void *h, double vol) { return static_cast<VortexState *>(h)->integrateZ(vol); }
int ref_vs_save(void *h, const char *path) { return static_cast<VortexState *>(h)->save(path) ? 1 : 0; }

// Picture::save (Output/Picture.cpp:81-137) of a width x height picture whose
// colours are rgba[4*(y*width + x) + c].
int ref_picture_save(unsigned width, unsigned height, const double *rgba, const char *path, int middle) {
    Picture pic(width, height);
    for (unsigned y = 0; y < height; y++)
        for (unsigned x = 0; x < width; x++) {
            const double *c = rgba + 4 * ((size_t)y * width + x);
            pic.setColor(x, y, Vector4(c[0], c[1], c[2], c[3]));
        }
    return pic.save(path, middle != 0) ? 1 : 0;
}

// saveVolumeMap (GPUGrossPitaevskii_BCC.cu:38-67) restated over the
// reference's own Text class, which formats the numbers; that function sits
// in a CUDA source file that cannot be compiled here.
int ref_save_volume_map(const char *path, const unsigned short *vol, unsigned xsize, unsigned ysize, unsigned zsize,
                        double hx, double hy, double hz) {
    Text rawpath;
    rawpath << path << ".raw";
    std::ofstream fs(rawpath.str().c_str(), std::ios_base::binary | std::ios::trunc);
    if (fs.fail()) return 0;
    fs.write((const char *)vol, 2 * xsize * ysize * zsize);
    fs.close();
    Text text;
    text << "ObjectType              = Image" << std::endl;
    text << "NDims                   = 3" << std::endl;
    text << "BinaryData              = True" << std::endl;
    text << "CompressedData          = False" << std::endl;
    text << "BinaryDataByteOrderMSB  = False" << std::endl;
    text << "TransformMatrix         = 1 0 0 0 1 0 0 0 1" << std::endl;
    text << "Offset                  = " << -0.5 * xsize * hx << " " << -0.5 * ysize * hy << " " << -0.5 * zsize * hz
         << std::endl;
    text << "CenterOfRotation        = 0 0 0" << std::endl;
    text << "DimSize                 = " << xsize << " " << ysize << " " << zsize << std::endl;
    text << "ElementSpacing          = " << hx << " " << hy << " " << hz << std::endl;
    text << "ElementNumberOfChannels = 1" << std::endl;
    text << "ElementType             = MET_USHORT" << std::endl;
    text << "ElementDataFile         = " << rawpath.str() << std::endl;
    return text.save(path) ? 1 : 0;
}

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
