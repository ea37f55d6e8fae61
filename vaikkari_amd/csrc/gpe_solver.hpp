// gpe_solver.hpp -- the opaque gpe_solver of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "gpe_internal.hpp"
#include "gpe_kernels.hpp"

namespace gpe {
Layout make_layout(uint32_t nx, uint32_t ny, uint32_t nz, size_t elem_bytes);
Trap make_trap(const double origin[3], double b, double ratio);
// Host [k][j][i][l] complex double <-> device class-row layout of one field
// (T = double or float), synchronous on stream st.
template <typename T> gpe_status upload_field(void *dev, const double *h, const Layout &L, hipStream_t st);
template <typename T> gpe_status download_field(const void *dev, double *h, const Layout &L, hipStream_t st);
}  // namespace gpe

struct gpe_solver {
    gpe_config cfg{};
    gpe::Layout L{};
    gpe::Trap trap{};
    gpe::Coefs co{};
    int kernel = GPE_KERNEL_GATHER;
    int tile = 0;
    bool tile_default = true;  // the family's default tile (0 at creation or in gpe_set_variant)
    int device = 0;
    // field[ie] holds the even time level, field[io] the odd one (or the
    // imaginary-time scratch buffer).  The fused kernel writes the step's
    // result into the other two buffers and swaps.
    void *field[4] = {nullptr, nullptr, nullptr, nullptr};
    int ie = 0, io = 1;
    void *pot = nullptr;           // V*dt in the field layout (GPE_POT_ARRAY)
    std::vector<double> V_host;    // unscaled V, host order (re-scaled on dt change)
    double *scal = nullptr;        // [0] last N, [1] read scale, [4..7] observables
    double *partials = nullptr;    // per-block partial sums
    long long npartials = 0;
    void *out_buf = nullptr;       // device staging of pictures / volume maps (grown on demand)
    size_t out_bytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t steps_done = 0;
    bool odd_valid = false;

    ~gpe_solver();
    gpe_status init(const gpe_config &c);
    gpe_status alloc_field(int f);
    void set_coefs();
    size_t elem_bytes() const;
    size_t nodes() const;
    size_t device_bytes() const;
    template <typename T> gpe_status upload(int which, const double *h);
    template <typename T> gpe_status download(int which, double *h);
    gpe_status upload_potential();
    template <typename T, int FORM> gpe_status enqueue_steps(uint64_t nsteps, uint64_t *launches);
    gpe_status step(uint64_t nsteps, uint64_t *launches);
    void imag_variant(int &k, int &t) const;
    template <typename T> gpe_status observables_t(int which, double out[4]);
    template <typename T, int FORM> gpe_status observe_step(int tile, int even_tile);
};
