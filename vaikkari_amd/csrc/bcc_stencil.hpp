// bcc_stencil.hpp -- the 4-neighbour stencil of the BCC dual mesh, usable
// from host and device code.  bcc_face(l, f) is a constexpr switch so that
// fully unrolled kernels resolve every neighbour to a register or to a
// fixed address offset at compile time.
#pragma once
#include <hip/hip_runtime.h>

namespace gpe {

struct Face {
    int dx, dy, dz, m;
};

// Neighbour f (0..3) of class l (0..11): the node of class m in the cell at
// offset (dx,dy,dz).  Every pair is at distance b (|q_m + 8*d - q_l|^2 = 8
// in units of b/sqrt(8)); the set is fixed by the geometry.  The order of
// the four faces per class is the reference kernel's summation order
// (koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu:93-140), kept so that the
// FP64 neighbour sum ((n0+n1)+n2)+n3 associates identically.
__host__ __device__ constexpr Face bcc_face(int l, int f) {
    switch (l * 4 + f) {
        case 0: return {0, 0, 0, 9};
        case 1: return {0, 0, 0, 10};
        case 2: return {-1, 0, 1, 2};
        case 3: return {-1, 0, 0, 3};
        case 4: return {0, 0, 0, 2};
        case 5: return {0, 0, 0, 3};
        case 6: return {1, -1, 0, 5};
        case 7: return {0, -1, 0, 8};
        case 8: return {0, 0, 0, 1};
        case 9: return {0, 0, 0, 4};
        case 10: return {1, 0, -1, 0};
        case 11: return {0, 0, -1, 11};
        case 12: return {0, 0, 0, 1};
        case 13: return {0, 0, 0, 4};
        case 14: return {1, 0, 0, 0};
        case 15: return {0, 0, 0, 11};
        case 16: return {0, 0, 0, 2};
        case 17: return {0, 0, 0, 3};
        case 18: return {1, 0, 0, 5};
        case 19: return {0, 0, 0, 8};
        case 20: return {0, 0, 0, 6};
        case 21: return {0, 0, 0, 7};
        case 22: return {-1, 1, 0, 1};
        case 23: return {-1, 0, 0, 4};
        case 24: return {0, 0, 0, 5};
        case 25: return {0, 0, 0, 8};
        case 26: return {0, 1, -1, 9};
        case 27: return {0, 0, -1, 10};
        case 28: return {0, 0, 0, 5};
        case 29: return {0, 0, 0, 8};
        case 30: return {0, 1, 0, 9};
        case 31: return {0, 0, 0, 10};
        case 32: return {0, 0, 0, 6};
        case 33: return {0, 0, 0, 7};
        case 34: return {0, 1, 0, 1};
        case 35: return {0, 0, 0, 4};
        case 36: return {0, 0, 0, 0};
        case 37: return {0, 0, 0, 11};
        case 38: return {0, -1, 1, 6};
        case 39: return {0, -1, 0, 7};
        case 40: return {0, 0, 0, 0};
        case 41: return {0, 0, 0, 11};
        case 42: return {0, 0, 1, 6};
        case 43: return {0, 0, 0, 7};
        case 44: return {0, 0, 0, 9};
        case 45: return {0, 0, 0, 10};
        case 46: return {0, 0, 1, 2};
        case 47: return {0, 0, 0, 3};
    }
    return {0, 0, 0, -1};
}

__host__ __device__ constexpr bool face_in_cell(Face e) { return e.dx == 0 && e.dy == 0 && e.dz == 0; }

}  // namespace gpe
