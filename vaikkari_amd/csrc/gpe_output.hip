// gpe_output.hip -- volume maps and slice pictures of a solver's even field,
// reduced on the device, the writers of the reference's output files
// (MetaImage volume, 24-bit BMP), and integrateInTime with the reference's
// SAVE_PICTURE / SAVE_VOLUME output.
//
// The reference copies the full padded field back each iteration and loops
// over it on the host (GPUGrossPitaevskii_BCC.cu:424-472).  A volume map
// reads each interior cell's 12 class values once (192 B/cell in FP64,
// 96 B in FP32; HBM-bound) and writes 2 bytes; a slice reads one class of
// one z layer.  Only those results cross PCIe.
#include <hip/hip_runtime.h>

#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/vaikkari_gpe_output.h"
#include "gpe_hip_check.hpp"
#include "gpe_kernels.hpp"
#include "gpe_solver.hpp"

namespace gpe {
namespace {

constexpr int kOutTX = 64, kOutTY = 4;  // threads per block: 64 along x, 4 rows

// min(65535, trunc(v)), NaN -> 0 (GPUGrossPitaevskii_BCC.cu:461).
__device__ __forceinline__ uint16_t clamp_u16(double v) {
    if (v > 65535.0) return 65535;
    if (v >= 0.0) return (uint16_t)v;
    return 0;
}

// One colour channel as Picture::save stores it (Output/Picture.cpp:123-126):
// 0 below 1 (and for NaN), 255 above 255, else truncated.
__device__ __forceinline__ uint8_t channel_byte(double c) {
    if (c > 255.0) return 255;
    if (c >= 1.0) return (uint8_t)(int)c;
    return 0;
}

// Element index of padded cell (px, py, pz), class 0, or -1 in the halo.
__device__ __forceinline__ long long interior_index(const Layout &L, int px, int py, int pz) {
    if (px < 1 || px > L.nx || py < 1 || py > L.ny || pz < 1 || pz > L.nz) return -1;
    return L.origin + (long long)(pz - 1) * L.sz + (long long)(py - 1) * L.sy + (px - 1);
}

// Volume map: blockIdx.z = padded z layer, (64 x 4) threads per block over
// padded (x, y).  Sums are FP64 in class order, the products squared in FP64
// after widening (the host's arithmetic on the downloaded field).
template <typename T, bool CLASS0>
__global__ void __launch_bounds__(kOutTX *kOutTY)
    k_volume_map(const cplx_t<T> *__restrict__ f, Layout L, double unit, uint16_t *__restrict__ out) {
#pragma clang fp contract(off)
    const int wx = L.nx + 2, wy = L.ny + 2;
    const int px = blockIdx.x * kOutTX + threadIdx.x, py = blockIdx.y * kOutTY + threadIdx.y, pz = blockIdx.z;
    if (px >= wx || py >= wy) return;
    const long long e = interior_index(L, px, py, pz);
    double s = 0.0;
    if (e >= 0) {
        if (CLASS0) {
            const cplx_t<T> v = f[e];
            const double x = v.x, y = v.y;
            for (int l = 0; l < GPE_CLASSES; l++) s += x * x + y * y;
        } else {
#pragma unroll
            for (int l = 0; l < GPE_CLASSES; l++) {
                const cplx_t<T> v = f[e + l * L.xp];
                const double x = v.x, y = v.y;
                s += x * x + y * y;
            }
        }
    }
    out[((size_t)pz * wy + py) * wx + px] = clamp_u16(s * unit);
}

// Slice picture of padded z layer pz: colour gain*(Re, |psi|, Im) of class 0
// (GPUGrossPitaevskii_BCC.cu:433-436), bytes as Picture::save (blue, green, red).
template <typename T>
__global__ void __launch_bounds__(kOutTX *kOutTY)
    k_slice_picture(const cplx_t<T> *__restrict__ f, Layout L, int pz, double gain, int middle,
                    uint8_t *__restrict__ out) {
#pragma clang fp contract(off)
    const int wx = L.nx + 2, wy = L.ny + 2;
    const int px = blockIdx.x * kOutTX + threadIdx.x, py = blockIdx.y * kOutTY + threadIdx.y;
    if (px >= wx || py >= wy) return;
    const long long e = interior_index(L, px, py, pz);
    double re = 0.0, im = 0.0;
    if (e >= 0) {
        const cplx_t<T> v = f[e];
        re = v.x;
        im = v.y;
    }
    const double nrm = sqrt(re * re + im * im);
    const double rgb[3] = {gain * re, gain * nrm, gain * im};
    uint8_t b[3];
    for (int c = 0; c < 3; c++) b[c] = channel_byte(middle ? 128.0 * rgb[c] + 128.0 : 256.0 * rgb[c]);
    uint8_t *o = out + 3 * ((size_t)py * wx + px);
    o[0] = b[2];
    o[1] = b[1];
    o[2] = b[0];
}

gpe_status ensure_out(gpe_solver *s, size_t bytes) {
    if (s->out_bytes >= bytes) return GPE_OK;
    if (s->out_buf) {
        GPE_HIP_TRY(hipStreamSynchronize(s->stream));
        GPE_HIP_TRY(hipFree(s->out_buf));
        s->out_buf = nullptr;
        s->out_bytes = 0;
    }
    if (hipMalloc(&s->out_buf, bytes) != hipSuccess) {
        (void)hipGetLastError();
        s->out_buf = nullptr;
        set_error("output staging: cannot allocate %zu bytes", bytes);
        return GPE_ENOMEM;
    }
    s->out_bytes = bytes;
    return GPE_OK;
}

dim3 out_grid(const Layout &L, int layers) {
    return dim3((unsigned)((L.nx + 2 + kOutTX - 1) / kOutTX), (unsigned)((L.ny + 2 + kOutTY - 1) / kOutTY),
                (unsigned)layers);
}

}  // namespace
}  // namespace gpe

using namespace gpe;

extern "C" {

gpe_status gpe_volume_map(gpe_solver *s, double unit, uint32_t flags, uint16_t *out) {
    if (!s || !out || !std::isfinite(unit) || unit < 0.0 || flags > GPE_VOLUME_REFERENCE_CLASS0) {
        set_error("gpe_volume_map: need a solver, an output array, finite unit >= 0 and a known flag");
        return GPE_EINVAL;
    }
    const Layout &L = s->L;
    const size_t cells = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    gpe_status st = ensure_out(s, cells * sizeof(uint16_t));
    if (st != GPE_OK) return st;
    auto *dout = static_cast<uint16_t *>(s->out_buf);
    const dim3 grid = out_grid(L, L.nz + 2), block(kOutTX, kOutTY);
    const void *f = s->field[s->ie];
    const bool c0 = flags == GPE_VOLUME_REFERENCE_CLASS0;
    if (s->cfg.precision == GPE_FP64) {
        auto *fd = static_cast<const cplx_t<double> *>(f);
        if (c0) hipLaunchKernelGGL((k_volume_map<double, true>), grid, block, 0, s->stream, fd, L, unit, dout);
        else hipLaunchKernelGGL((k_volume_map<double, false>), grid, block, 0, s->stream, fd, L, unit, dout);
    } else {
        auto *ff = static_cast<const cplx_t<float> *>(f);
        if (c0) hipLaunchKernelGGL((k_volume_map<float, true>), grid, block, 0, s->stream, ff, L, unit, dout);
        else hipLaunchKernelGGL((k_volume_map<float, false>), grid, block, 0, s->stream, ff, L, unit, dout);
    }
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipMemcpyAsync(out, dout, cells * sizeof(uint16_t), hipMemcpyDeviceToHost, s->stream));
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    return GPE_OK;
}

gpe_status gpe_slice_picture(gpe_solver *s, uint32_t k, double gain, int32_t middle, uint8_t *bgr) {
    if (!s || !bgr || k > (uint32_t)s->L.nz + 1 || !std::isfinite(gain)) {
        set_error("gpe_slice_picture: need a solver, an output array, a finite gain and 0 <= k <= nz+1");
        return GPE_EINVAL;
    }
    const Layout &L = s->L;
    const size_t bytes = (size_t)3 * (L.nx + 2) * (L.ny + 2);
    gpe_status st = ensure_out(s, bytes);
    if (st != GPE_OK) return st;
    auto *dout = static_cast<uint8_t *>(s->out_buf);
    const dim3 grid = out_grid(L, 1), block(kOutTX, kOutTY);
    const int mid = middle ? 1 : 0;
    if (s->cfg.precision == GPE_FP64)
        hipLaunchKernelGGL((k_slice_picture<double>), grid, block, 0, s->stream,
                           static_cast<const cplx_t<double> *>(s->field[s->ie]), L, (int)k, gain, mid, dout);
    else
        hipLaunchKernelGGL((k_slice_picture<float>), grid, block, 0, s->stream,
                           static_cast<const cplx_t<float> *>(s->field[s->ie]), L, (int)k, gain, mid, dout);
    GPE_HIP_TRY(hipGetLastError());
    GPE_HIP_TRY(hipMemcpyAsync(bgr, dout, bytes, hipMemcpyDeviceToHost, s->stream));
    GPE_HIP_TRY(hipStreamSynchronize(s->stream));
    return GPE_OK;
}

// The MetaImage header saveVolumeMap writes (GPUGrossPitaevskii_BCC.cu:51-64):
// "Key = value" lines with the key left-aligned in 24 columns, numbers in the
// default stream format, the data file named by its full path.
gpe_status gpe_write_volume_map(const char *path, const uint16_t *vol, const uint32_t dims[3],
                                const double spacing[3]) {
    if (!path || !vol || !dims || !spacing) {
        set_error("gpe_write_volume_map: null argument");
        return GPE_EINVAL;
    }
    const std::string raw = std::string(path) + ".raw";
    {
        std::ofstream fs(raw, std::ios::binary | std::ios::trunc);
        if (!fs) {
            set_error("gpe_write_volume_map: cannot open %s", raw.c_str());
            return GPE_EIO;
        }
        fs.write((const char *)vol, (std::streamsize)(2 * (size_t)dims[0] * dims[1] * dims[2]));
        if (!fs) {
            set_error("gpe_write_volume_map: write failed on %s", raw.c_str());
            return GPE_EIO;
        }
    }
    std::ostringstream off, dim, sp;
    off << -0.5 * dims[0] * spacing[0] << " " << -0.5 * dims[1] * spacing[1] << " " << -0.5 * dims[2] * spacing[2];
    dim << dims[0] << " " << dims[1] << " " << dims[2];
    sp << spacing[0] << " " << spacing[1] << " " << spacing[2];
    const std::pair<const char *, std::string> lines[] = {
        {"ObjectType", "Image"},
        {"NDims", "3"},
        {"BinaryData", "True"},
        {"CompressedData", "False"},
        {"BinaryDataByteOrderMSB", "False"},
        {"TransformMatrix", "1 0 0 0 1 0 0 0 1"},
        {"Offset", off.str()},
        {"CenterOfRotation", "0 0 0"},
        {"DimSize", dim.str()},
        {"ElementSpacing", sp.str()},
        {"ElementNumberOfChannels", "1"},
        {"ElementType", "MET_USHORT"},
        {"ElementDataFile", raw},
    };
    std::ostringstream text;
    for (const auto &kv : lines) text << std::left << std::setw(24) << kv.first << "= " << kv.second << "\n";
    std::ofstream hs(path, std::ios::trunc);
    if (!hs) {
        set_error("gpe_write_volume_map: cannot open %s", path);
        return GPE_EIO;
    }
    hs << text.str();
    if (!hs) {
        set_error("gpe_write_volume_map: write failed on %s", path);
        return GPE_EIO;
    }
    return GPE_OK;
}

// 24-bit BMP as Picture::save writes it (Output/Picture.cpp:81-137): a
// 14-byte file header and a 40-byte info header, whose size fields count 4
// bytes per pixel, then the rows, row 0 first, each padded with zeros to a
// multiple of 4 bytes.
gpe_status gpe_write_bmp(const char *path, uint32_t width, uint32_t height, const uint8_t *bgr) {
    if (!path || (!bgr && width && height)) {
        set_error("gpe_write_bmp: null argument");
        return GPE_EINVAL;
    }
    std::ofstream fs(path, std::ios::binary | std::ios::trunc);
    if (!fs) {
        set_error("gpe_write_bmp: cannot open %s", path);
        return GPE_EIO;
    }
    auto u32 = [&fs](uint32_t v) { fs.write((const char *)&v, 4); };
    auto u16 = [&fs](uint16_t v) { fs.write((const char *)&v, 2); };
    const uint32_t offs = 14 + 40, data = 4u * (width * height), size = offs + data;
    fs.write("BM", 2);
    u32(size);
    u32(0);
    u32(offs);
    u32(40);
    u32(width);
    u32(height);
    u16(1);
    u16(24);
    u32(0);
    u32(data);
    u32(0);
    u32(0);
    u32(0);
    u32(0);
    const uint32_t pad = 3 - ((3 * width + 3) % 4);
    const char zeros[4] = {0, 0, 0, 0};
    for (uint32_t y = 0; y < height; y++) {
        fs.write((const char *)bgr + (size_t)3 * width * y, (std::streamsize)3 * width);
        if (pad) fs.write(zeros, pad);
    }
    if (!fs) {
        set_error("gpe_write_bmp: write failed on %s", path);
        return GPE_EIO;
    }
    return GPE_OK;
}

}  // extern "C"
