// checkpoint.cpp -- save and resume a solver's run (include/vaikkari_gpe.h,
// gpe_checkpoint_save / gpe_checkpoint_load).
//
// The reference saves only the VortexState that starts a run
// (koodit/GrossPitaevskii/VortexState.cpp:74-96); a run itself cannot be
// interrupted and resumed.  A checkpoint here holds what the next step reads:
// the configuration in force (mode and dt as set last, the kernel family and
// tile in use), the steps taken, both time levels and a streamed potential.
// FP32 fields are stored as doubles, which hold every float exactly, so a
// resumed solver continues the run bit for bit in either precision.
//
// File layout (little-endian):
//   "VGPECKP1", u32 version (1),
//   u32 nx, ny, nz, precision, potential, mode, kernel, tile (0: the defaults),
//   f64 block_scale, dt, g, origin[3], trap_ratio,
//   u64 steps_done, u32 has_odd, u32 has_potential,
//   f64 even[2N], f64 odd[2N] (if has_odd), f64 V[N] (if has_potential),
//   u64 FNV-1a 64 of every byte before it.
// N = 12 nx ny nz nodes; arrays in the host order of the C ABI.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gpe_internal.hpp"
#include "gpe_solver.hpp"

namespace {

constexpr char kMagic[8] = {'V', 'G', 'P', 'E', 'C', 'K', 'P', '1'};
constexpr uint32_t kVersion = 1;

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        for (size_t i = 0; i < n; i++) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    }
};

struct Writer {
    FILE *f;
    Fnv sum;
    bool ok = true;
    void put(const void *p, size_t n) {
        if (ok && n && std::fwrite(p, 1, n, f) != n) ok = false;
        sum.add(p, n);
    }
    template <typename T> void val(T v) { put(&v, sizeof v); }
};

struct Reader {
    const unsigned char *p, *end;
    Fnv sum;
    bool get(void *dst, size_t n) {
        if ((size_t)(end - p) < n) return false;
        std::memcpy(dst, p, n);
        sum.add(p, n);
        p += n;
        return true;
    }
    template <typename T> bool val(T &v) { return get(&v, sizeof v); }
};

}  // namespace

extern "C" {

gpe_status gpe_checkpoint_save(gpe_solver *s, const char *path) {
    if (!s || !path) {
        gpe::set_error("gpe_checkpoint_save: null argument");
        return GPE_EINVAL;
    }
    const size_t nodes = s->nodes();
    const bool has_odd = s->cfg.mode == GPE_REAL_TIME && s->odd_valid;
    const bool has_pot = s->cfg.potential == GPE_POT_ARRAY;
    if (has_pot && s->V_host.size() != nodes) {
        gpe::set_error("gpe_checkpoint_save: the solver streams a potential but none was set");
        return GPE_EINVAL;
    }
    std::vector<double> even(2 * nodes), odd(has_odd ? 2 * nodes : 0);
    gpe_status st = gpe_get_state(s, even.data(), has_odd ? odd.data() : nullptr);
    if (st != GPE_OK) return st;
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        gpe::set_error("gpe_checkpoint_save: cannot open %s for writing", path);
        return GPE_EIO;
    }
    Writer w{f};
    w.put(kMagic, sizeof kMagic);
    w.val(kVersion);
    // tile 0 when the solver runs its family's defaults (real and imaginary time may differ)
    const uint32_t tile = s->tile_default ? 0u : (uint32_t)s->tile + 1;
    const uint32_t u[8] = {s->cfg.nx,        s->cfg.ny,   s->cfg.nz,           s->cfg.precision,
                           s->cfg.potential, s->cfg.mode, (uint32_t)s->kernel, tile};
    w.put(u, sizeof u);
    const double d[7] = {s->cfg.block_scale, s->cfg.dt,        s->cfg.g,          s->cfg.origin[0],
                         s->cfg.origin[1],   s->cfg.origin[2], s->cfg.trap_ratio};
    w.put(d, sizeof d);
    w.val((uint64_t)s->steps_done);
    w.val((uint32_t)(has_odd ? 1 : 0));
    w.val((uint32_t)(has_pot ? 1 : 0));
    w.put(even.data(), even.size() * sizeof(double));
    if (has_odd) w.put(odd.data(), odd.size() * sizeof(double));
    if (has_pot) w.put(s->V_host.data(), nodes * sizeof(double));
    const uint64_t sum = w.sum.h;
    if (w.ok && std::fwrite(&sum, sizeof sum, 1, f) != 1) w.ok = false;
    if (std::fclose(f) != 0) w.ok = false;
    if (!w.ok) {
        gpe::set_error("gpe_checkpoint_save: writing %s failed", path);
        return GPE_EIO;
    }
    return GPE_OK;
}

gpe_status gpe_checkpoint_load(const char *path, gpe_solver **out) {
    if (!path || !out) {
        gpe::set_error("gpe_checkpoint_load: null argument");
        return GPE_EINVAL;
    }
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        gpe::set_error("gpe_checkpoint_load: cannot open %s", path);
        return GPE_EIO;
    }
    std::vector<unsigned char> bytes;
    {
        unsigned char buf[1 << 16];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + n);
        const bool err = std::ferror(f) != 0;
        std::fclose(f);
        if (err) {
            gpe::set_error("gpe_checkpoint_load: reading %s failed", path);
            return GPE_EIO;
        }
    }
    auto bad = [&](const char *why) {
        gpe::set_error("gpe_checkpoint_load: %s is not a valid checkpoint (%s)", path, why);
        return GPE_EIO;
    };
    Reader r{bytes.data(), bytes.data() + bytes.size()};
    char magic[8];
    uint32_t version = 0, u[8], has_odd = 0, has_pot = 0;
    double d[7];
    uint64_t steps = 0;
    if (!r.get(magic, sizeof magic) || std::memcmp(magic, kMagic, sizeof magic) != 0) return bad("magic");
    if (!r.val(version) || version != kVersion) return bad("version");
    if (!r.get(u, sizeof u) || !r.get(d, sizeof d) || !r.val(steps) || !r.val(has_odd) || !r.val(has_pot))
        return bad("short header");
    gpe_config cfg{};
    cfg.nx = u[0];
    cfg.ny = u[1];
    cfg.nz = u[2];
    cfg.precision = u[3];
    cfg.potential = u[4];
    cfg.mode = u[5];
    cfg.kernel = u[6];
    cfg.tile = u[7];
    cfg.block_scale = d[0];
    cfg.dt = d[1];
    cfg.g = d[2];
    cfg.origin[0] = d[3];
    cfg.origin[1] = d[4];
    cfg.origin[2] = d[5];
    cfg.trap_ratio = d[6];
    if (cfg.nx == 0 || cfg.ny == 0 || cfg.nz == 0 || cfg.nx > 4096 || cfg.ny > 4096 || cfg.nz > 4096 ||
        has_odd > 1 || has_pot > 1 || (has_pot != 0) != (cfg.potential == GPE_POT_ARRAY))
        return bad("header fields");
    const size_t nodes = (size_t)GPE_CLASSES * cfg.nx * cfg.ny * cfg.nz;
    const size_t want = (2 + (has_odd ? 2 : 0) + (has_pot ? 1 : 0)) * nodes * sizeof(double) + sizeof(uint64_t);
    if ((size_t)(r.end - r.p) != want) return bad("size");
    std::vector<double> even(2 * nodes), odd(has_odd ? 2 * nodes : 0), V(has_pot ? nodes : 0);
    r.get(even.data(), even.size() * sizeof(double));
    if (has_odd) r.get(odd.data(), odd.size() * sizeof(double));
    if (has_pot) r.get(V.data(), V.size() * sizeof(double));
    uint64_t sum = 0;
    const uint64_t expect = r.sum.h;
    std::memcpy(&sum, r.p, sizeof sum);
    if (sum != expect) return bad("checksum");

    // a streamed potential is given after creation, as a caller would
    gpe_solver *s = nullptr;
    gpe_status st = gpe_create(&cfg, &s);
    if (st != GPE_OK) return st;
    if (has_pot) st = gpe_set_potential(s, V.data());
    if (st == GPE_OK) st = gpe_set_state(s, even.data(), has_odd ? odd.data() : nullptr);
    if (st != GPE_OK) {
        gpe_destroy(s);
        return st;
    }
    s->steps_done = steps;
    *out = s;
    return GPE_OK;
}

gpe_status gpe_get_config(const gpe_solver *s, gpe_config *out) {
    if (!s || !out) {
        gpe::set_error("gpe_get_config: null argument");
        return GPE_EINVAL;
    }
    *out = s->cfg;
    out->kernel = (uint32_t)s->kernel;
    out->tile = (uint32_t)s->tile + 1;
    return GPE_OK;
}

}  // extern "C"
