// lx_internal.hpp -- device data layout and kernel parameter blocks.
//
// Device layout (one simulation):
//   * Each field component is one array of nz planes x ny rows x px columns,
//     x fastest (px = nx rounded up to 64 elements).  The reference's C order
//     has z fastest (lib/field/field.h:54); x-fastest lets every wavefront
//     read 64 consecutive x cells while the kernels march along z.
//   * All six components share that shape; a component's staggered axis uses
//     n-1 entries (lib/field/vfield.cc:75-84) and the unused tail stays zero.
//   * H is stored scaled, Hs = eta0*H, and lengths are measured in dx, so every
//     update coefficient is O(1) (see lx_api.cpp for the derivation).
//   * Materials: one uint8 id per node, all E components at a storage index
//     share it; cb[id] = c*dt/(dx*eps_r) lives in a 256-entry table.
//   * CPML psi lives only in the boundary slabs: per axis a, the cells with
//     index < slo[a] or >= shi[a], with four FP32 arrays per axis
//     (two H and two E components that carry a derivative along a).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>

namespace lx {

constexpr int kBX = 64;   // threads along x per block (one wavefront)
constexpr int kBY = 4;    // rows per block

// One TF/SF correction box: the cells (of one component, one split curl, one
// face of the TF box) whose curl difference gets  dir * g(node)  added.
struct TfBox {
    int lo[3];       // inclusive index box; empty when lo[0] > hi[0]
    int hi[3];
    float off[3];    // node coordinate = index + off (cells)
    double dir;      // signed direction factor in device units (0: inactive)
};

// Boxes of one half step: [component][curl 0/1][face lo/hi].
struct TfSet {
    TfBox box[3][2][2];
    int any;
};

// Incident-field provider (see lx_api.cpp: analytic = reference Gaussian,
// aux = FP64 1D line updated on the device).
struct IncParams {
    int mode;            // LX_INC_ANALYTIC / LX_INC_AUX1D
    int planar;
    double dx, dy, dz, kx, ky, kz;
    double init_dx;      // off_pw * dx
    double cdt;          // c * dt
    double inv_sigma;    // 1 / pw_sigma
    int axis, x0, L;
    const double *e;     // line E at integer nodes   [L]
    const double *h;     // line eta0*H at half nodes [L-1]
};

template <typename T>
struct Fields {
    T *E[3];
    T *H[3];
};

// Coefficient tables per axis (length n_a, FP32, computed in FP64 on host):
//   kh = ch*(dx/d_a)/kappa_h, bh = b_h, chh = ch*c_h*(dx/d_a)   (H update)
//   ke = (dx/d_a)/kappa_e,    be = b_e, ce = c_e*(dx/d_a)       (E update)
struct Tables {
    const float *kh[3], *bh[3], *chh[3];
    const float *ke[3], *be[3], *ce[3];
};

struct Slabs {
    int lo[3], hi[3], w[3];   // slab [0,lo) U [hi,n); thickness w = lo + n - hi
    float *psiH[3][2];        // [axis][0: curl1 of comp (a+2)%3, 1: curl2 of comp (a+1)%3]
    float *psiE[3][2];
};

template <typename T>
struct Grid {
    int nx, ny, nz, px;
    long long plane;          // px * ny
    int kz;                   // z cells per block
    Fields<T> f;
    const uint8_t *mat;
    const float *cb;          // [256]
    Tables tab;
    Slabs slab;
    int pml;
    int *counter;             // ping-pong step counter [2]
};

// Launchers (lx_kernels.hip).
template <typename T>
void launch_h3(const Grid<T> &g, const TfSet &tf, const IncParams &inc, hipStream_t s);
template <typename T>
void launch_e3(const Grid<T> &g, const TfSet &tf, const IncParams &inc, hipStream_t s);
template <typename T>
void launch_h2(const Grid<T> &g, const TfSet &tf, const IncParams &inc, hipStream_t s);
template <typename T>
void launch_e2(const Grid<T> &g, const TfSet &tf, const IncParams &inc, hipStream_t s);

struct AuxLine {
    double *e, *h;
    int L;
    double S;             // c*dt/d_axis
    int wave;
    double c, dt, lambda0, fwhm, delay;
};
void launch_aux_h(const AuxLine &a, hipStream_t s);
void launch_aux_e(const AuxLine &a, const int *counter, hipStream_t s);

struct PointSrc {
    int kind, comp, wave;
    long long idx;        // device linear index of the source cell
    double amp;           // soft: amplitude; hard: int(amp)
    int omega;            // hard: int(omega)
    double dt, c, lambda0, fwhm, delay;
};
template <typename T>
void launch_point_source(const Grid<T> &g, const PointSrc &ps, hipStream_t s);

struct ProbeSet {
    int n;
    const long long *idx;   // device linear index per probe
    const int *comp;        // component per probe
    double *rec;            // [cap * n]
    int cap;
    double hscale;          // 1/eta0
};
template <typename T>
void launch_probes(const Grid<T> &g, const ProbeSet &p, hipStream_t s);

// Gather one component into FP64 reference layout (C order, z fastest).
template <typename T>
void launch_gather(const Grid<T> &g, int comp, int ex, int ey, int ez, double scale, double *dst,
                   int planar, hipStream_t s);

// FP64 monitor: out[0] = max |E|^2 (reference index range), out[1] = energy sum,
// out[2] = non-finite count.
template <typename T>
void launch_monitor(const Grid<T> &g, int planar, double *out, hipStream_t s);

void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
void launch_zero_bytes(void *p, size_t n, hipStream_t s);

}  // namespace lx
