// lx_util.hpp -- error state and small helpers shared by the host sources.
#pragma once
#include <cstdarg>
#include <cstdio>

#include "../../include/luxum.h"

namespace lx {

// Thread-local last error message (lx_last_error).
char *error_buffer();

inline lx_status fail(lx_status st, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(error_buffer(), 512, fmt, ap);
    va_end(ap);
    return st;
}

// Yee stagger (x,y,z) per component (lib/field/vfield.cc:75-84).
constexpr int kStag[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

// Host extents of a component; planar runs hold Ey (Nx,1,Nz), Hx (Nx,1,Nz-1),
// Hz (Nx-1,1,Nz) (vfield.cc:62-72).  Returns false for absent components.
inline bool comp_extent(const int N[3], int planar, int comp, int ext[3]) {
    if (planar) {
        ext[1] = 1;
        if (comp == LX_EY) { ext[0] = N[0]; ext[2] = N[2]; return true; }
        if (comp == LX_HX) { ext[0] = N[0]; ext[2] = N[2] - 1; return true; }
        if (comp == LX_HZ) { ext[0] = N[0] - 1; ext[2] = N[2]; return true; }
        ext[0] = ext[1] = ext[2] = 0;
        return false;
    }
    for (int a = 0; a < 3; ++a) ext[a] = N[a] - kStag[comp][a];
    return true;
}

}  // namespace lx
