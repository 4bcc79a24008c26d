// lx_internal.hpp -- device data layout and kernel parameter blocks.
//
// Device layout (one simulation):
//   * The six field components live in one allocation, component c at
//     F + c*fstride (Ex,Ey,Ez,Hx,Hy,Hz).  Each is nz planes x ny rows x px
//     columns, x fastest (px = nx rounded up to 64 elements).  The
//     reference's C order has z fastest (lib/field/field.h:54); x-fastest
//     lets a wavefront read 64 consecutive x cells while the kernels march
//     along z.
//   * A component's staggered axis uses n-1 entries (lib/field/vfield.cc:75-84);
//     the unused tail stays zero.
//   * H is stored scaled, Hs = eta0*H, and lengths are measured in dx, so every
//     update coefficient is O(1) (lx_api.cpp has the derivation).
//   * Materials: one uint8 id per node shared by the E components at that
//     storage index; each E component c has its own 256-entry table,
//     cb[c*256 + id] = c*dt/(dx*eps_r,c), so that the three components may
//     see different permittivities as the reference's per-component Cb
//     arrays do (maxwell_three_dim_pml.cc:109-128).  Planar runs use the Ey
//     table (c = 1).
//   * CPML psi lives only in the boundary slabs: per axis a, the cells with
//     index < slo[a] or >= shi[a]; four FP32 arrays per axis (psi of the two H
//     and the two E components that carry a derivative along a).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>

namespace lx {

// Hardware sends block b to XCD b % 8; give XCD x a contiguous range of the
// n tiles (or jobs), so that neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, l = b >> 3;
    return x * q + min(x, r) + l;
}

// Wavefront shifts (DPP wave_shl1 / wave_shr1): lane l receives lane l+1
// (next) or l-1 (prev); lane 63 (next) and lane 0 (prev) receive 0.
__device__ __forceinline__ float wave_next(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float wave_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}

constexpr int kBX = 64;      // threads along x per block (one wavefront)
constexpr int kBY = 4;       // rows per block
constexpr int kMaxTf = 12;   // TF/SF correction boxes per half step

// Coefficient tables: tab[(a*6 + kind)*tn + i], kind:
//   0 kh = ch*(dx/d_a)/kappa_h   1 bh = b_h   2 chh = ch*c_h*(dx/d_a)   (H update)
//   3 ke = (dx/d_a)/kappa_e      4 be = b_e   5 ce = c_e*(dx/d_a)       (E update)
enum { TK_KH = 0, TK_BH = 1, TK_CH = 2, TK_KE = 3, TK_BE = 4, TK_CE = 5 };

// One TF/SF correction box: the cells (of one component, one split curl, one
// face of the TF box) whose curl difference gets  dir * g(node)  added.
struct TfBox {
    int lo[3];       // inclusive index box
    int hi[3];
    float off[3];    // node coordinate = index + off (cells)
    int comp;        // target component 0..2 (of E or H)
    int curl;        // 0: first split curl, 1: second
    int fa;          // axis normal to the face
    int start;       // first flattened cell index of this box
    double dir;      // signed direction factor in device units
};

struct TfList {
    int n;
    int total;             // face cells over all boxes
    TfBox b[kMaxTf];       // host copy
    const TfBox *dev;      // device copy read by the kernels
};

// Incident-field provider.
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

// FP64 aux-1D incident line, advanced by an extra block of the bulk kernels.
struct AuxLine {
    double *e, *h;
    double *e2, *h2;     // second set: the launch between two fused steps reads one set and writes the other
    int L;
    double S;            // c*dt/d_axis
    int wave;
    double c, dt, lambda0, fwhm, delay;
};

// Probes: E probes are recorded by the H launch of the next step (E is final
// and read-only then), H probes by the E launch of the same step; the last
// step's E probes by the sequence-end kernel.
struct ProbeArgs {
    int n;
    const int *comp;         // device [n]
    const long long *idx;    // device [n], linear index into the component
    double *rec;             // device [cap * n]
    int cap;
    double hscale;           // 1/eta0
};

template <typename T>
struct Grid {
    int nx, ny, nz, px;
    long long plane;          // px * ny
    long long fstride;        // elements between components
    T *F;
    const uint8_t *mat;
    const float *cb;          // [3][256]: Ex, Ey, Ez tables
    const float *tab;
    int tn;                   // table row length
    float *psi;
    long long scells[3];      // psi cells per array on each axis
    int slo[3], shi[3], sw[3];
    int pml;
    int kz;                   // z planes per block (3D) / rows per block (2D)
    int kzs;                  // fused 3D step: z planes per block of the CPML shell (0: kz)
    int ntx, nty, ntz, ntiles;
    int *counter;             // counter[0] = steps completed before this launch sequence
    T *Fn;                    // fused 3D step: fields written (F is read)
    float *psin;              // fused 3D step: psi written (psi is read)
    int rows;                 // fused 3D step: rows per block of the interior launch sharing their
                              // neighbours through LDS (0: one unit per wavefront, no sharing)
};

struct Launch {
    int toff;                 // step offset of this launch from counter[0]
    int aux;                  // advance the aux line in the extra block
    int probes;               // record probes in the extra block
    int eprobes;              // the H launch records the E probes of the previous step (else the
                              // point-source kernel does, before it changes its cell)
};

// Bulk half-step kernel: half 0 = H update, 1 = E update.
template <typename T>
void launch_bulk(const Grid<T> &g, int planar, int half, hipStream_t s);

// TF/SF face corrections of one half step (+ aux line + probes in an extra block).
template <typename T>
void launch_tfsf(const Grid<T> &g, const TfList &tf, const IncParams &inc, const AuxLine &aux, const ProbeArgs &pr,
                 Launch l, int half, hipStream_t s);

// Fused 3D timestep (H and E half steps, F/psi read, Fn/psin written), nw
// wavefronts per block, vec (1, 2, or 4 with FP16 storage) x cells per lane;
// vec needs the padded row pitch to hold every 64*vec-cell chunk (px >= nx
// rounded up to 64*vec).  Bulk update only: the H TF/SF corrections are
// applied to the read buffer beforehand (launch_tfsf_hpre), the E ones to the
// written buffer afterwards (launch_tfsf).  Two parts: part 0 is the whole
// grid without CPML, else the interior box (rows and planes outside the y and
// z slabs); part 1 is the CPML shell around it (nothing without CPML).
template <typename T>
void launch_step3(const Grid<T> &g, int nw, int vec, int part, hipStream_t s);
// Rows and planes of the part-0 box of a CPML grid.
void step3_core_extent(int ny, int nz, int kz, const int slo[3], const int shi[3], int *rows, int *planes);

// The units of one fused-step launch, all x chunks of each.  Mode 0: the
// rows [jlo, jhi) x the z blocks [tlo, thi) of depth kz (the interior box).
// Mode 1: every cell outside that box, in z blocks of depth kz: the rows
// outside [jlo, jhi) over the whole z range (ntz blocks), then the rows inside
// over [0, zlo) (nlo blocks) and [zhi, nz).
struct UnitMap {
    int mode, jlo, jhi, tlo, thi, ntz, kz, zlo, zhi, nlo;
};
// Row jr and planes [k0, k1) of the rest-th (row, z block) pair of a map.
__host__ __device__ inline void unit_span(const UnitMap &um, int ny, int nz, int rest, int &jr, int &k0, int &k1) {
    if (um.mode == 0) {
        const int nr = um.jhi - um.jlo;
        jr = um.jlo + rest % nr;
        k0 = (um.tlo + rest / nr) * um.kz;
        k1 = min(k0 + um.kz, nz);
        return;
    }
    const int nro = um.jlo + (ny - um.jhi);    // rows outside the box: every z block
    if (rest < nro * um.ntz) {
        const int r = rest % nro;
        jr = r < um.jlo ? r : um.jhi + (r - um.jlo);
        k0 = (rest / nro) * um.kz;
        k1 = min(k0 + um.kz, nz);
        return;
    }
    rest -= nro * um.ntz;                      // rows inside: z blocks below and above the box
    const int nr = um.jhi - um.jlo, t = rest / nr;
    jr = um.jlo + rest % nr;
    if (t < um.nlo) {
        k0 = t * um.kz;
        k1 = min(k0 + um.kz, um.zlo);
    } else {
        k0 = um.zhi + (t - um.nlo) * um.kz;
        k1 = min(k0 + um.kz, nz);
    }
}
// Unit maps and unit counts of the two parts of a grid's fused step (part 1
// is empty without CPML) at vec cells per lane: interior box in z blocks of
// kz planes, shell in blocks of kzs (0: kz).
void step3_unit_maps(int nx, int ny, int nz, int kz, int kzs, int pml, const int slo[3], const int shi[3], int vec,
                     UnitMap um[2], int nunits[2]);

// H TF/SF corrections of the step applied to H of the read buffer before the
// fused step: H -= s*kh*delta per face cell (the fused step then computes
// (H - s*kh*delta) - curl).  Needs every H face outside the CPML slabs (no
// psi term); an extra block advances the aux line's H phase and records the
// E probes of the previous step.
template <typename T>
void launch_tfsf_hpre(const Grid<T> &g, const TfList &tf, const IncParams &inc, const AuxLine &aux,
                      const ProbeArgs &pr, Launch l, hipStream_t s);

// Fused 3D step, between steps t and t+1 (g.F = the buffer step t wrote,
// Fprev = the one it read): the E corrections of step t and the H
// corrections of step t+1 in one launch, the aux line from set ain to set
// aout, the H probes of step t and (l.eprobes) the E probes of step t-1.
// last: no H corrections (the sequence's final step).  Requires no H probe
// on an H face cell.
// Blocks after the face-cell blocks of such a launch: one thread per aux
// node (L nodes, 0 without an aux line), at least one when there are probes.
inline int between_extra_blocks(int L, bool probes) {
    const int n = (L + 255) / 256;
    return n > 0 ? n : (probes ? 1 : 0);
}
template <typename T>
void launch_tfsf_between(const Grid<T> &g, const T *Fprev, const TfList &tfE, const TfList &tfH, const IncParams &inc,
                         const AuxLine &ain, const AuxLine &aout, const ProbeArgs &pr, Launch l, int last,
                         hipStream_t s);

// End of a launch sequence of n steps: records the E probes of the last step
// and advances counter[0] by n.
template <typename T>
void launch_finish(const Grid<T> &g, const ProbeArgs &pr, int n, hipStream_t s);

struct PointSrc {
    int kind, comp, wave;
    long long idx;        // device linear index of the source cell
    double amp;           // soft: amplitude; hard: int(amp)
    int omega;            // hard: int(omega)
    double dt, c, lambda0, fwhm, delay;
};
// Point source of step counter[0] + toff on E of g.F; first records the E
// probes of the previous step (E as that step left it).
template <typename T>
void launch_point_source(const Grid<T> &g, const PointSrc &ps, const ProbeArgs &pr, int toff, hipStream_t s);

// Gather one component into FP64 reference layout (C order, z fastest).
template <typename T>
void launch_gather(const Grid<T> &g, int comp, int ex, int ey, int ez, double scale, double *dst, hipStream_t s);

// |E|^2 at equal storage indices, FP64, into the Ez-shaped host layout
// (ex, ey, ez = Ez extents), zero outside the reference's loop range.
template <typename T>
void launch_gather_e2(const Grid<T> &g, int ex, int ey, int ez, double scale, double *dst, hipStream_t s);

// FP64 monitor: out[0] = max |E|^2 (reference index range), out[1] = sum of
// eps_r*E^2 + Hs^2 (0 with e_only), out[2] = non-finite count (of the E
// components only with e_only); out holds 4 + 3*kMonBlocks
// doubles (the per-block partial triples follow the result).  With res the
// result triple goes there instead (host-mapped memory: no copy follows).
constexpr int kMonBlocks = 2048;
template <typename T>
void launch_monitor(const Grid<T> &g, int planar, double cb0, double *out, bool e_only, hipStream_t s,
                    double *res = nullptr);

void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
// material ids, host C order (i, j, k) -> device layout (k, j, x)
void launch_ids_to_device(const uint8_t *src, uint8_t *dst, int nx, int ny, int nz, int px, long long plane,
                          hipStream_t s);
void launch_zero_bytes(void *p, size_t n, hipStream_t s);

}  // namespace lx
