// ses3d_hip.hip -- MI355X (gfx950) time-stepper for the SES3D spherical-section
// spectral-element solver (lpd = 4), behind the C-ABI of include/ses3d_hip.h.
//
// Data layout in HBM (per section):
//   element-local fields  slot e*128 + nd    e = i + NX*(j + NY*k), nd = l + 5*(m + 5*n)
//                                            (125 nodes in a 512-byte aligned slot per element;
//                                            the element force in node-block order, fperm)
//                         components stored in pairs (two floats per point side by side) so a
//                         lane moves two components with one 8-byte access
//   per-event state       one contiguous block per element (StateLayout): split stresses,
//                         memory variables and displacement gradient in pair planes
//   global fields         v[blocked]         4x4x4 node blocks: node (gx,gy,gz) lives in
//                                            block (gx>>2, gy>>2, gz>>2) at (gz&3)*16+(gy&3)*4+(gx&3);
//                                            (vx, vy) as one pair plane, vz alone
// Per step (SOURCE/ses3d_evolution.f90:7-373):
//   k_element    (recording steps) its first workgroups sample the seismograms from the
//                pre-update velocity (ses3d_output.f90:475-495), one per receiver; the
//                others: gather v -> 9 velocity gradients (5-term GLL contractions through LDS)
//                -> spherical terms -> displacement-gradient update -> absorbing-layer terms
//                -> stress rates (elastic or 3-mechanism anelastic) -> split-stress update
//                -> weak-form divergence (LDS) + moment source -> point forces / adjoint
//                sources -> element force; memory-variable update, stress taper,
//                forward-field store (optionally at a lower polynomial degree) and (adjoint)
//                kernel accumulation fused
//   k_node       gather the element forces of the <=8 elements sharing a node in the
//                reference's assembly order, divide by the mass, velocity update, taper
// All float32 arithmetic follows the reference's left-to-right order and is compiled with
// -ffp-contract=off, so results are bit-identical to the reference's float32 build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ses3d_hip.h"
#include "setup.hpp"

using ses3d::NN;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail(SES3D_HIP_EDEVICE, std::string(#expr ": ") + hipGetErrorString(_e));  \
    } while (0)

#ifndef SES3D_EPB
#define SES3D_EPB 1
#endif
// minimum waves per SIMD the element kernels are compiled for: 5 fits them in
// 96 VGPRs without spilling, except the field-storing kernels (mode 1), which
// are compiled for 4 (at 5 they spill two registers and run 0.5-0.8% slower)
#ifndef SES3D_WAVES_PER_EU
#define SES3D_WAVES_PER_EU 5
#endif
#define WAVES_MIN(MODE, EPB) ((MODE) == 1 ? 4 : (SES3D_WAVES_PER_EU ? SES3D_WAVES_PER_EU : 1))
// Launch schedule: elements per element-kernel workgroup (1, 2 or 4; 128
// lanes each, 125 active), lanes per node-kernel workgroup (128, 256 or 512:
// 2, 4 or 8 node blocks of 64) and the XCD-aware workgroup order.  Every
// schedule computes bit-identical results; the autotuner picks by timing.
constexpr int DEFAULT_EPB = SES3D_EPB;
constexpr int TPE = 128;                 // threads per element (125 active)
#ifndef SES3D_EST
#define SES3D_EST 128
#endif
#ifndef SES3D_CSPAD
#define SES3D_CSPAD 4224
#endif
// Element-local arrays: each element's 125 values in a 512-byte aligned slot of
// 128; consecutive components' arrays are offset by a further 16.5 KiB so the
// many per-point streams of a workgroup do not start at the same HBM channel
// offset.  (Measured on MI355X: node kernel 7% faster than with 125-value
// slots, element kernel within 1%.)
constexpr int EST = SES3D_EST;           // element stride of element-local arrays (125 used)
constexpr int CSPAD = SES3D_CSPAD;       // floats added to each component array
constexpr int DEFAULT_NODE_THREADS = 256;
#ifndef SES3D_NODE_MAX_THREADS
#define SES3D_NODE_MAX_THREADS 512
#endif
constexpr int NODE_MAX_THREADS = SES3D_NODE_MAX_THREADS;

// component slots (same order for e, d, split stresses)
enum { EXX = 0, EYY, EZZ, EXY, EYX, EXZ, EZX, EYZ, EZY };

struct MeshDev {
    int NX, NY, NZ;            // elements per axis
    int NBX, NBY, NBZ;         // node blocks per axis (NX+1, ...)
    long long NE, NP, NGB;     // elements, element-local slots (NE*EST), padded global nodes
    int KX, KY, KZ;
    int is_diss;
    float dt, Jac, sum_D_p;
    float tau_p[3], Lk[3];
    float tab[7 * 25 + 40];    // D1 D2 D3 WDL DUM1 DUM2 DUM3, then the stored-field
                               // degree tables HLC (5x4) and HLD (4x5)
    const float *lm;   // 2*NP: (lambda, mu) pairs
    const float *A, *B, *C, *Q, *rhoinv, *MM;
    const float *Z, *S, *COT;
    const unsigned short *IX, *IY, *IZ;
    const float *LUT;
    const unsigned char *PMLX, *PMLY, *PMLZ;   // per element index: inside an absorbing layer
    // quotients by the relaxation times and by 3 through a reciprocal and two
    // FMAs, enabled only for divisors whose every float32 numerator was
    // checked at init to give the IEEE quotient bit for bit
    float rtau_p[3], r3;
    int fast_tau, fast_3;
    // stored forward fields: degree hdeg (4: unchanged), hnc = (hdeg+1)^3
    // values per element and field, hslot floats per sampled step
    int hdeg, hnc;
    long long hslot;
    // kernel accumulators in float32 (the reference's type) instead of float64
    int kf32;
    int nt;
};

// One point force or adjoint source (SOURCE/ses3d_miscellaneous.f90:29-92).
struct ForceDev {
    int comp_mask;        // bit c: subtract from component c
    const float *w;       // 125 weights
    const float *stf;     // 3*nt: component c at stf[c*nt + it-1]
};

// A receiver (SOURCE/ses3d_output.f90:456-497): the event's global velocity
// it samples, its element, its 125 interpolation weights (index l + 5m + 25n,
// precomputed on the host) and its 3 x nt seismogram (component-major).
struct RecDev {
    const float *v;     // (vx, vy) pair plane, vz at 2*ngb
    const int *itp;     // the section's step number
    long long ngb;
    int nbx, nby;
    int ex, ey, ez;
    int nt;
    const float *w;
    float *out;
};

struct EventDev {
    float *v;          // 3*NGB: (vx, vy) pair plane, then vz at 2*NGB
    float *st;         // NE state blocks of StateLayout<is_diss>::ESB floats: each element's
                       // split stresses, memory variables and displacement gradient
    float *f;          // 3*NP  element force: (fx, fy) pair plane, then fz at 2*NP;
                       // nodes in fperm order within each element slot
    int src_elem;      // moment-tensor source element or -1
    const float *src_md;   // 9*125 MOM_c*delta
    const float *so;       // nt source time function
    float *hist;           // sampled forward fields, 9*NP per slot: pair planes (vx,vy)
                           // (vz,exx) (eyy,ezz) (exy,exz), eyz at 8*NP (or degree < 4 values)
    const int *slot;       // [it] -> slot or -1 (nt+1 entries)
    double *grad;          // 5*NP  pair planes (rho,cp) (csh,csv), Q_mu at 4*NP; with
                           // float32 accumulators the same layout in floats
    float sdt;             // samp_ad*dt
    // point forces / adjoint sources: per element {first, count} into
    // `forces` (the element's sources in their reference order), or nullptr
    // when the event has none
    const int2 *force_elem;
    const ForceDev *forces;
};

// Loads and stores of the streamed per-point state (touched once per step).
// SES3D_NT_LOAD / SES3D_NT_STORE select non-temporal cache hints for them.
// Pointers to HBM are cast to the global address space so loads and stores
// compile to global_* instructions (counted by vmcnt only): LDS accesses then
// need not wait for outstanding HBM loads.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float cgfloat;
#define GLOB(p) ((gfloat *)(p))
#define CGLOB(p) ((cgfloat *)(p))

#ifndef SES3D_NT_LOAD
#define SES3D_NT_LOAD 0
#endif
#ifndef SES3D_NT_STORE
#define SES3D_NT_STORE 0
#endif
template <typename T>
__device__ __forceinline__ T ldg(const T *p) {
    return *((__attribute__((address_space(1))) const T *)p);
}

__device__ __forceinline__ float ldnt(const float *p) {
#if SES3D_NT_LOAD
    return __builtin_nontemporal_load(CGLOB(p));
#else
    return *CGLOB(p);
#endif
}
// element-local access: uniform array base (scalar registers) + 32-bit byte
// offset of the lane, so the address needs no per-lane 64-bit arithmetic
// (every element-local array is below 4 GiB, checked on the host).  The
// element kernel's state streams (d, s, M, coefficients, sampled fields) are
// read and written once per step, so ldo/sto and ldo2/sto2 are non-temporal:
// they leave the caches to the data that is reused (the gathered velocity,
// the tables, the element force the node kernel reads back).  Measured on the
// 96x96x48 workload: element kernel -7 to -10%.
typedef __attribute__((address_space(1))) const char cgchar;
typedef __attribute__((address_space(1))) char gchar;
template <int NT>
__device__ __forceinline__ float ldo(const float *base, unsigned boff) {
    if (NT) return __builtin_nontemporal_load((cgfloat *)((cgchar *)base + boff));
    return *(cgfloat *)((cgchar *)base + boff);
}
template <int NT>
__device__ __forceinline__ void sto(float *base, unsigned boff, float v) {
    if (NT) __builtin_nontemporal_store(v, (gfloat *)((gchar *)base + boff));
    else *(gfloat *)((gchar *)base + boff) = v;
}

// Paired components: two components of one point side by side, so a lane
// moves them with one 8-byte access (a wavefront then moves 512 contiguous
// bytes per instruction with half the memory instructions).  A pair plane
// holds 2*NP floats, point p's pair at floats 2p, 2p+1.
//   d, s (9 components): planes 0..3 = (EXX,EYY) (EZZ,EXY) (EXZ,EYZ) (EYX,EZX),
//                        EZY alone at float 8*NP (stride 1);
//   M (18 components, index c*3+k): plane q = (2q, 2q+1).
// Outside the absorbing layers only planes 0..2 of s are touched.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) d2v gd2v;
typedef __attribute__((address_space(1))) const f2v cgf2v;
typedef __attribute__((address_space(1))) f2v gf2v;
template <int NT>
__device__ __forceinline__ f2v ldo2(const float *base, unsigned boff) {
    if (NT) return __builtin_nontemporal_load((cgf2v *)((cgchar *)base + boff));
    return *(cgf2v *)((cgchar *)base + boff);
}
template <int NT>
__device__ __forceinline__ void sto2(float *base, unsigned boff, float a, float b) {
    f2v v;
    v.x = a;
    v.y = b;
    if (NT) __builtin_nontemporal_store(v, (gf2v *)((gchar *)base + boff));
    else *(gf2v *)((gchar *)base + boff) = v;
}
// the element force: written by the element kernel and read back by the node
// kernel right after, so through the caches (plain stores and loads)
__device__ __forceinline__ void stc2(float *base, unsigned boff, float a, float b) {
    f2v v;
    v.x = a;
    v.y = b;
    *(gf2v *)((gchar *)base + boff) = v;
}
__device__ __forceinline__ void stc(float *base, unsigned boff, float v) { *(gfloat *)((gchar *)base + boff) = v; }
__device__ __forceinline__ f2v ldc2(const float *base, unsigned boff) { return *(cgf2v *)((cgchar *)base + boff); }
__device__ __forceinline__ float ldc(const float *base, unsigned boff) { return *(cgfloat *)((cgchar *)base + boff); }
// Position of node (l,m,n) in an element's slot of the element force f:
// first the 4x4x4 nodes with l,m,n < 4 (the node block whose lanes own them in
// the node kernel) in block-lane order, then the l=4, m=4 and n=4 faces, the
// three edges and the corner, each in the order its neighbouring block's lanes
// read it.  A node-kernel wavefront then reads each sharer's contributions as
// one contiguous run instead of scattered over the 125-node slot.
__host__ __device__ constexpr int fperm(int l, int m, int n) {
    return (l < 4 && m < 4 && n < 4) ? l + 4 * (m + 4 * n)
         : (l == 4 && m < 4 && n < 4) ? 64 + m + 4 * n
         : (l < 4 && m == 4 && n < 4) ? 80 + l + 4 * n
         : (l < 4 && m < 4 && n == 4) ? 96 + l + 4 * m
         : (l == 4 && m == 4 && n < 4) ? 112 + n
         : (l == 4 && m < 4 && n == 4) ? 116 + m
         : (l < 4 && m == 4 && n == 4) ? 120 + l
         : 124;
}

// Per-event state: one contiguous block per element holding its 128-node
// slots of the split stresses, memory variables and displacement gradient,
// so a workgroup streams its element's state as one run of memory (fewer,
// longer DRAM bursts than one array per component).  Floats from the block
// start: stress pair planes (xx,yy) (zz,xy) (xz,yz) first, then the 9 memory
// variable pair planes (anelastic), which together are all an element outside
// the absorbing layers touches, then the stress pair (yx,zx) and zy, then the
// displacement gradient laid out as the stresses.
struct StateLayout {
    static constexpr int PS = 2 * TPE;   // floats of one pair plane of one element
    int m_off, s_out, d_off, esb;
    __host__ __device__ constexpr StateLayout(int diss)
        : m_off(3 * PS), s_out(3 * PS + (diss ? 9 * PS : 0)), d_off(s_out + PS + TPE), esb(d_off + 4 * PS + TPE) {}
};
// offset in a 9-component (stress or displacement-gradient) group of a block,
// and the stride between nodes: pairs (xx,yy) (zz,xy) (xz,yz) at 0, PS, 2PS
// (inner part), (yx,zx) at `outer`, zy at outer + PS
__host__ __device__ constexpr int c9_block_offset(int c, int outer) {
    return c == EXX ? 0 : c == EYY ? 1 : c == EZZ ? StateLayout::PS : c == EXY ? StateLayout::PS + 1
         : c == EXZ ? 2 * StateLayout::PS : c == EYZ ? 2 * StateLayout::PS + 1 : c == EYX ? outer
         : c == EZX ? outer + 1 : outer + StateLayout::PS;
}
__host__ __device__ constexpr int c9_stride(int c) { return c == EZY ? 1 : 2; }
// inner: the six components every element touches (three pair planes from b)
template <int NT>
__device__ __forceinline__ void ld9_inner(const float *b, unsigned o2, float *x) {
    f2v t = ldo2<NT>(b, o2);
    x[EXX] = t.x; x[EYY] = t.y;
    t = ldo2<NT>(b + StateLayout::PS, o2);
    x[EZZ] = t.x; x[EXY] = t.y;
    t = ldo2<NT>(b + 2 * StateLayout::PS, o2);
    x[EXZ] = t.x; x[EYZ] = t.y;
}
// outer: the second members of the split pairs (pair plane at b, EZY after it)
template <int NT>
__device__ __forceinline__ void ld9_outer(const float *b, unsigned o2, unsigned o, float *x) {
    const f2v t = ldo2<NT>(b, o2);
    x[EYX] = t.x; x[EZX] = t.y;
    x[EZY] = ldo<NT>(b + StateLayout::PS, o);
}
template <int NT>
__device__ __forceinline__ void st9_inner(float *b, unsigned o2, const float *x) {
    sto2<NT>(b, o2, x[EXX], x[EYY]);
    sto2<NT>(b + StateLayout::PS, o2, x[EZZ], x[EXY]);
    sto2<NT>(b + 2 * StateLayout::PS, o2, x[EXZ], x[EYZ]);
}
template <int NT>
__device__ __forceinline__ void st9_outer(float *b, unsigned o2, unsigned o, const float *x) {
    sto2<NT>(b, o2, x[EYX], x[EZX]);
    sto<NT>(b + StateLayout::PS, o, x[EZY]);
}

__device__ __forceinline__ void stnt(float *p, float v) {
#if SES3D_NT_STORE
    __builtin_nontemporal_store(v, GLOB(p));
#else
    *GLOB(p) = v;
#endif
}

// a/b from y = RN(1/b): q = RN(a y), exact remainder, one correction step
// (Markstein).  Below |a| = 2^-96 the remainder can underflow and the IEEE
// division is used.  Enabled only for divisors for which this function was
// checked against a / b for every float32 numerator.
__device__ __forceinline__ float div_recip(float a, float b, float y) {
    if (!(__builtin_fabsf(a) >= 0x1p-96f)) return a / b;
    const float q = a * y;
    const float r = __builtin_fmaf(-q, b, a);
    return __builtin_fmaf(r, y, q);
}

__device__ __forceinline__ long long gidx(int gx, int gy, int gz, int NBX, int NBY) {
    return ((((long long)(gz >> 2) * NBY + (gy >> 2)) * NBX + (gx >> 2)) << 6) +
           (((gz & 3) << 4) | ((gy & 3) << 2) | (gx & 3));
}

// Seismogram sample of one receiver (SOURCE/ses3d_output.f90:477-491) by one
// workgroup of >= 128 lanes: lane t forms the product of term t of the
// reference's sum (i slowest, k fastest) for the three components, then one
// lane per component adds the 125 products in that order from 0, so every
// product and every partial sum is rounded as in the sequential loop while the
// 125 gathers are issued at once.  `prod`: 3 x 128 floats of LDS.
__device__ __forceinline__ void receiver_sample(const RecDev &R, float *prod) {
    const int t = threadIdx.x;
    if (t < 125) {
        const int i = t / 25, j = (t / 5) % 5, k = t % 5;
        const long long g = gidx(4 * R.ex + i, 4 * R.ey + j, 4 * R.ez + k, R.nbx, R.nby);
        const f2v vxy = *(cgf2v *)(R.v + 2 * g);
        const float w = R.w[i + 5 * (j + 5 * k)];
        prod[t] = vxy.x * w;
        prod[128 + t] = vxy.y * w;
        prod[256 + t] = R.v[2 * R.ngb + g] * w;
    }
    __syncthreads();
    if (t < 3) {
        const int it = *R.itp;
        float acc = 0.f;
        for (int q = 0; q < 125; ++q) acc = acc + prod[t * 128 + q];
        R.out[t * R.nt + it - 1] = acc;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so neighbouring elements share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    int q = nb >> 3, r = nb & 7, x = bid & 7, loc = bid >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + loc;
}

// Point forces and adjoint sources of one element (SOURCE/ses3d_miscellaneous.f90:
// 29-92, called at ses3d_evolution.f90:276): the element force at node nd
// less w_nd * stf(it) of each source, in the sources' order.  Sources are
// read in chunks of FCH (independent loads, one round trip), the
// subtractions then run in order.
template <typename FT>
__device__ __forceinline__ void subtract_forces(FT *__restrict__ F, const int count, const int nd, const int nt,
                                                const int it, float &fx, float &fy, float &fz) {
    constexpr int FCH = 8;
    for (int s0 = 0; s0 < count; s0 += FCH) {
        float w[FCH], sx[FCH], sy[FCH], sz[FCH];
        int mask[FCH];
#pragma unroll
        for (int j = 0; j < FCH; ++j) {
            mask[j] = 0;
            w[j] = sx[j] = sy[j] = sz[j] = 0.0f;
            if (s0 + j < count) {
                const ForceDev Fj = F[s0 + j];
                mask[j] = Fj.comp_mask;
                w[j] = Fj.w[nd];
                if (Fj.comp_mask & 1) sx[j] = Fj.stf[0 * nt + it - 1];
                if (Fj.comp_mask & 2) sy[j] = Fj.stf[1 * nt + it - 1];
                if (Fj.comp_mask & 4) sz[j] = Fj.stf[2 * nt + it - 1];
            }
        }
#pragma unroll
        for (int j = 0; j < FCH; ++j) {
            if (mask[j] & 1) fx = fx - w[j] * sx[j];
            if (mask[j] & 2) fy = fy - w[j] * sy[j];
            if (mask[j] & 4) fz = fz - w[j] * sz[j];
        }
    }
}

// ---------------------------------------------------------------------------
// element kernel
// MODE 0: forward; 1: forward, store sampled fields when slot[it] >= 0;
// 2: adjoint, accumulate kernels when slot[it] >= 0.
// ANISO 0: the model has A = B = C = 0 everywhere (checked on the host), so
// their terms, exact zeros in the reference's sums, are not formed.
//
// Elements outside every absorbing layer have prof_x = prof_y = prof_z = 0 and
// taper 1 at all their nodes.  There the split stresses of each pair
// (sxy,syx), (sxz,szx), (syz,szy) obey the same update from the same zero
// start, so they are bit-identical and only one of each pair is stored; and in
// forward mode (MODE 0) the displacement gradient feeds nothing but the
// absorbing terms, so it is integrated only inside the layers.
// ---------------------------------------------------------------------------
// One element-kernel workgroup: EPB elements of one event.  `blk` is the
// workgroup's index among the section's blocks_per_event * n_events tiles.
// MT / ET: MeshDev / EventDev, plain or in the constant address space.
template <int DISS, int MODE, int ANISO, int EPB, int NT, typename MT, typename ET>
__device__ __forceinline__ void element_step(MT &Ms, ET *__restrict__ evs, const int it, const int blk,
                                             const int blocks_per_event, const RecDev *rec) {
    const int ev = blk / blocks_per_event;
    const int eb = blk - ev * blocks_per_event;
    auto &E = evs[ev];
    const int t = threadIdx.x;
    // a wavefront (64 lanes) never straddles two elements (128 lanes each)
    const int sub = __builtin_amdgcn_readfirstlane(t >> 7), nd = t & (TPE - 1);
    const int NE = (int)Ms.NE;
    const size_t NP = (size_t)Ms.NP;
    const int e = eb * EPB + sub;
    // Lanes 125..127 of an element compute copies of node 124 and read and
    // write only the unused tail of the element's 128-value slots (never read
    // for a real node), so no branch depends on the lane: a wavefront works
    // iff its element exists.
    const bool act = e < NE;
    const int ndc = nd < 125 ? nd : 124;

    __shared__ float tabs[7 * 25 + 40];
    // stored fields of degree < 4: the step's 9 fields of each element (MODE 1,
    // to be compressed) or its stored values (MODE 2, to be interpolated)
    __shared__ float hv[MODE ? EPB : 1][MODE ? 9 * TPE : 1];
    __shared__ float sv[EPB][3][TPE];
    __shared__ float ss[EPB][9][TPE];
    if (rec) {
        // a receiver workgroup of a recording step (uniform), in the stress buffer's LDS
        receiver_sample(*rec, &ss[0][0][0]);
        return;
    }
    for (int k = t; k < 7 * 25 + 40; k += EPB * TPE) tabs[k] = Ms.tab[k];
    const bool hcomp = MODE != 0 && Ms.hdeg < 4;   // uniform
    const float *D1 = tabs, *D2 = tabs + 25, *D3 = tabs + 50, *WDL = tabs + 75;
    const float *DUM1 = tabs + 100, *DUM2 = tabs + 125, *DUM3 = tabs + 150;

    const int l = ndc % 5, m = (ndc / 5) % 5, n = ndc / 25;
    const int NX = Ms.NX, NY = Ms.NY;
    const int ei = e % NX, ej = (e / NX) % NY, ek = e / (NX * NY);
    const int p = e * EST + nd;
    const unsigned pb = 4u * (unsigned)p;   // byte offset in every element-local array
    const unsigned pb2 = 2u * pb;           // byte offset in the pair planes
    // element in an absorbing layer? (uniform over the element's wavefronts)
    const bool pml = act && (ldg(Ms.PMLX + ei) | ldg(Ms.PMLY + ej) | ldg(Ms.PMLZ + ek));
    const bool keep_d = (MODE != 0) || pml;
    // the element's state block (uniform base) and the lane's offsets in its planes
    constexpr StateLayout SL(DISS);
    float *const sb = E.st + (size_t)(act ? e : 0) * SL.esb;
    const unsigned o = 4u * (unsigned)nd, o2 = 2u * o;

    float vx = 0.f, vy = 0.f, vz = 0.f;
    float d9[9], s9[9], Mv[18];
    float lam = 0.f, mu = 0.f, A = 0.f, B = 0.f, C = 0.f, QQ = 1.f;
    // The per-point state and coefficients.  In the plain forward step they are
    // issued after the first barrier, at the start of the gradient
    // contractions: fewer loads in flight per wavefront while the velocity is
    // gathered moves the state streams faster (element kernel -2.5% at
    // 96x96x48).  With the displacement gradient kept everywhere (modes 1 and
    // 2) the early issue is faster, and the order is the original one.
    auto load_coef = [&] {
        if (DISS) QQ = ldo<NT>(Ms.Q, pb);
        const f2v t = ldo2<NT>(Ms.lm, pb2);
        lam = t.x;
        mu = t.y;
    };
    auto load_d = [&] {
        if (keep_d) {
            ld9_inner<NT>(sb + SL.d_off, o2, d9);
            ld9_outer<NT>(sb + SL.d_off + 3 * StateLayout::PS, o2, o, d9);
        }
    };
    auto load_s = [&] {
        ld9_inner<NT>(sb, o2, s9);
        if (pml) ld9_outer<NT>(sb + SL.s_out, o2, o, s9);
    };
    auto load_m = [&] {
        if (DISS) {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const f2v t = ldo2<NT>(sb + SL.m_off + q * StateLayout::PS, o2);
                Mv[2 * q] = t.x;
                Mv[2 * q + 1] = t.y;
            }
        }
    };
    constexpr bool late_state = MODE == 0;
    if (act) {
        // gather v (SOURCE/ses3d_evolution.f90:27-51)
        const long long g = gidx(4 * ei + l, 4 * ej + m, 4 * ek + n, Ms.NBX, Ms.NBY);
        const f2v vxy = *(cgf2v *)(E.v + 2 * g);
        vx = vxy.x;
        vy = vxy.y;
        vz = ldnt(E.v + 2 * Ms.NGB + g);
        if (!late_state) {
            load_d();
            load_s();
            load_m();
            load_coef();
        }
        if (ANISO) {
            A = ldo<NT>(Ms.A, pb);
            B = ldo<NT>(Ms.B, pb);
            C = ldo<NT>(Ms.C, pb);
        }
        if (MODE == 2 && hcomp) {
            // the element's stored values of the paired forward step, [field][node]
            const int sl = ldg(E.slot + it);
            if (sl >= 0) {
                const int NC = Ms.hnc;
                const float *hs = E.hist + (size_t)sl * Ms.hslot + (size_t)e * NC;
                for (int t2 = nd; t2 < 9 * NC; t2 += TPE) {
                    const int c = t2 / NC;
                    hv[sub][t2] = ldg(hs + (size_t)c * NE * NC + (t2 - c * NC));
                }
            }
        }
    }
    sv[sub][0][nd] = vx;
    sv[sub][1][nd] = vy;
    sv[sub][2][nd] = vz;
    __syncthreads();

    const float dt = Ms.dt;
    if (act) {
        if (late_state) {
            load_coef();
            load_d();
            load_m();
            load_s();
        }
        // velocity gradients by 5-term contractions (:63-87); accumulate over q in order
        float ezz = 0.f, ezy = 0.f, ezx = 0.f, eyy = 0.f, eyx = 0.f, eyz = 0.f, exx = 0.f, exy = 0.f, exz = 0.f;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int qz = l + 5 * (m + 5 * q), qy = l + 5 * (q + 5 * n), qx = q + 5 * (m + 5 * n);
            const float c3 = D3[q * 5 + n], c2 = D2[q * 5 + m], c1 = D1[q * 5 + l];
            ezz = ezz + sv[sub][2][qz] * c3;
            ezy = ezy + sv[sub][1][qz] * c3;
            ezx = ezx + sv[sub][0][qz] * c3;
            eyy = eyy + sv[sub][1][qy] * c2;
            eyx = eyx + sv[sub][0][qy] * c2;
            eyz = eyz + sv[sub][2][qy] * c2;
            exx = exx + sv[sub][0][qx] * c1;
            exy = exy + sv[sub][1][qx] * c1;
            exz = exz + sv[sub][2][qx] * c1;
        }
        // spherical completion (:93-103)
        const float r = ldg(Ms.Z + ek * 5 + n), st = ldg(Ms.S + ei * 5 + l), ct = ldg(Ms.COT + ei * 5 + l);
        const float rst = r * st;
        ezz = -ezz;
        eyy = eyy / rst + (vz + vx * ct) / r;
        exx = (vz + exx) / r;
        ezx = -ezx;
        exz = (exz - vx) / r;
        ezy = -ezy;
        eyz = -vy / r + eyz / rst;
        exy = exy / r;
        eyx = -((vy * ct) / r) + eyx / rst;
        // displacement gradients (:109-117)
        float dxux = 0.f, dyuy = 0.f, dzuz = 0.f, dxuy = 0.f, dyux = 0.f, dxuz = 0.f, dzux = 0.f, dyuz = 0.f, dzuy = 0.f;
        if (keep_d) {
            dzuz = d9[EZZ] + dt * ezz;
            dyuy = d9[EYY] + dt * eyy;
            dxux = d9[EXX] + dt * exx;
            dzux = d9[EZX] + dt * ezx;
            dxuz = d9[EXZ] + dt * exz;
            dzuy = d9[EZY] + dt * ezy;
            dyuz = d9[EYZ] + dt * eyz;
            dxuy = d9[EXY] + dt * exy;
            dyux = d9[EYX] + dt * eyx;
            float *d = sb + SL.d_off;
            sto2<NT>(d, o2, dxux, dyuy);
            sto2<NT>(d + StateLayout::PS, o2, dzuz, dxuy);
            sto2<NT>(d + 2 * StateLayout::PS, o2, dxuz, dyuz);
            sto2<NT>(d + 3 * StateLayout::PS, o2, dyux, dzux);
            sto<NT>(d + 4 * StateLayout::PS, o, dzuy);
        }
        // absorbing-layer terms (:123-134); exact no-ops outside the layers
        float px = 0.f, py = 0.f, pz = 0.f, taper = 1.f;
        if (pml) {
            const unsigned ix = ldg(Ms.IX + ei * 5 + l), iy = ldg(Ms.IY + ej * 5 + m), iz = ldg(Ms.IZ + ek * 5 + n);
            const float *lut = Ms.LUT + (((size_t)iz * Ms.KY + iy) * Ms.KX + ix) * 5;
            px = ldg(lut); py = ldg(lut + 1); pz = ldg(lut + 2); taper = ldg(lut + 4);
            exx = exx + (pz + py) * dxux;
            eyy = eyy + (px + pz) * dyuy;
            ezz = ezz + (px + py) * dzuz;
            exy = exy + (py + pz) * dxuy;
            eyx = eyx + (px + pz) * dyux;
            exz = exz + (py + pz) * dxuz;
            ezx = ezx + (px + py) * dzux;
            eyz = eyz + (px + pz) * dyuz;
            ezy = ezy + (px + py) * dzuy;
        }
        exy = (exy + eyx) / 2.0f;
        exz = (exz + ezx) / 2.0f;
        eyz = (eyz + ezy) / 2.0f;
        // stress rates (:146-202)
        float sxx, syy, szz, sxy, sxz, syz;
        if (DISS) {
            const float tau = 1.0f / QQ;
            const float kappa = lam + (2.0f * mu) / 3.0f;
            const float mut = (1.0f + Ms.sum_D_p * tau) * mu;
            float Mdx = ((0.f + Mv[0]) + Mv[1]) + Mv[2];
            float Mdy = ((0.f + Mv[3]) + Mv[4]) + Mv[5];
            float Mdz = ((0.f + Mv[6]) + Mv[7]) + Mv[8];
            const float Md = (Mdx + Mdy) + Mdz;
            const float mt2 = (2.0f * mu) * tau;
            const float t1 = 2.0f * mut, t2 = mt2 * Md;
            const float d1 = Ms.fast_3 ? div_recip(t1, 3.0f, Ms.r3) : t1 / 3.0f;
            const float d2 = Ms.fast_3 ? div_recip(t2, 3.0f, Ms.r3) : t2 / 3.0f;
            const float L = (kappa - d1) * ((exx + eyy) + ezz) - d2;
            sxx = (L + (2.0f * mut) * exx) + mt2 * Mdx;
            syy = (L + (2.0f * mut) * eyy) + mt2 * Mdy;
            szz = (L + (2.0f * mut) * ezz) + mt2 * Mdz;
            if (ANISO) {
                sxx = (sxx + C * ezz) + A * (eyy + exx);
                syy = (syy + C * ezz) + A * (eyy + exx);
                szz = szz + C * (eyy + exx);
            }
            Mdx = ((0.f + Mv[9]) + Mv[10]) + Mv[11];
            Mdy = ((0.f + Mv[12]) + Mv[13]) + Mv[14];
            Mdz = ((0.f + Mv[15]) + Mv[16]) + Mv[17];
            sxy = (2.0f * mut) * exy + mt2 * Mdx;
            sxz = (2.0f * mut) * exz + mt2 * Mdy;
            syz = (2.0f * mut) * eyz + mt2 * Mdz;
            if (ANISO) {
                sxz = sxz + (2.0f * B) * exz;
                syz = syz + (2.0f * B) * eyz;
            }
            // memory variables (:334-360), from the post-absorbing, symmetrised rates
            const float ev6[6] = {exx, eyy, ezz, exy, exz, eyz};
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float Mo = Mv[c * 3 + k];
                    const float dM = dt * Mo;
                    const float q = Ms.fast_tau ? div_recip(dM, Ms.tau_p[k], Ms.rtau_p[k]) : dM / Ms.tau_p[k];
                    const float Mn = (Mo - q) - Ms.Lk[k] * ev6[c];
                    Mv[c * 3 + k] = Mn;
                    if ((c * 3 + k) & 1)
                        sto2<NT>(sb + SL.m_off + ((c * 3 + k) >> 1) * StateLayout::PS, o2, Mv[c * 3 + k - 1], Mn);
                }
        } else {
            const float L = lam * ((exx + eyy) + ezz);
            sxx = L + (2.0f * mu) * exx;
            syy = L + (2.0f * mu) * eyy;
            szz = L + (2.0f * mu) * ezz;
            if (ANISO) {
                sxx = (sxx + C * ezz) + A * (eyy + exx);
                syy = (syy + C * ezz) + A * (eyy + exx);
                szz = szz + C * (eyy + exx);
            }
            sxy = (2.0f * mu) * exy;
            sxz = (2.0f * mu) * exz;
            syz = (2.0f * mu) * eyz;
            if (ANISO) {
                sxz = sxz + (2.0f * B) * exz;
                syz = syz + (2.0f * B) * eyz;
            }
        }
        // split-stress integration (:208-219)
        if (pml) {
            s9[EXX] = s9[EXX] + dt * (sxx - px * s9[EXX]);
            s9[EYY] = s9[EYY] + dt * (syy - py * s9[EYY]);
            s9[EZZ] = s9[EZZ] + dt * (szz - pz * s9[EZZ]);
            s9[EXY] = s9[EXY] + dt * (sxy - px * s9[EXY]);
            s9[EYX] = s9[EYX] + dt * (sxy - py * s9[EYX]);
            s9[EXZ] = s9[EXZ] + dt * (sxz - px * s9[EXZ]);
            s9[EZX] = s9[EZX] + dt * (sxz - pz * s9[EZX]);
            s9[EYZ] = s9[EYZ] + dt * (syz - py * s9[EYZ]);
            s9[EZY] = s9[EZY] + dt * (syz - pz * s9[EZY]);
        } else {
            s9[EXX] = s9[EXX] + dt * sxx;
            s9[EYY] = s9[EYY] + dt * syy;
            s9[EZZ] = s9[EZZ] + dt * szz;
            s9[EXY] = s9[EXY] + dt * sxy;
            s9[EXZ] = s9[EXZ] + dt * sxz;
            s9[EYZ] = s9[EYZ] + dt * syz;
        }
        // moment-tensor stress glut (misc.f90:131-142, added at :251-264); the
        // paired glut components are equal, so the stored pair stays equal
        const int ncomp = pml ? 9 : 6;
        const int slot6[9] = {EXX, EYY, EZZ, EXY, EXZ, EYZ, EYX, EZX, EZY};
        if (e == E.src_elem) {
            const float so = ldg(E.so + it - 1);
#pragma unroll
            for (int k = 0; k < 9; ++k)
                if (k < ncomp) {
                    const int c = slot6[k];
                    ss[sub][c][nd] = s9[c] + (0.0f - ldg(E.src_md + c * 125 + ndc) * so);
                }
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k)
                if (k < ncomp) ss[sub][slot6[k]][nd] = s9[slot6[k]];
        }
        // stress taper (:368-371), 1 outside the layers; the divergence uses
        // the untapered stresses
        if (pml) {
#pragma unroll
            for (int c = 0; c < 9; ++c) s9[c] = s9[c] * taper;
            st9_inner<NT>(sb, o2, s9);
            st9_outer<NT>(sb + SL.s_out, o2, o, s9);
        } else {
            st9_inner<NT>(sb, o2, s9);
        }

        if (MODE == 1 && hcomp) {
            // fields to compress after the barrier (ses3d_main.f90:142-166)
            if (ldg(E.slot + it) >= 0) {
                float *hw = hv[sub];
                hw[0 * TPE + nd] = vx; hw[1 * TPE + nd] = vy; hw[2 * TPE + nd] = vz;
                hw[3 * TPE + nd] = dxux; hw[4 * TPE + nd] = dyuy; hw[5 * TPE + nd] = dzuz;
                hw[6 * TPE + nd] = (dxuy + dyux) / 2.0f;
                hw[7 * TPE + nd] = (dxuz + dzux) / 2.0f;
                hw[8 * TPE + nd] = (dyuz + dzuy) / 2.0f;
            }
        } else if (MODE == 1) {
            // forward-field store (ses3d_main.f90:142-166)
            const int sl = ldg(E.slot + it);
            if (sl >= 0) {
                float *h = E.hist + (size_t)sl * 9 * NP;
                sto2<NT>(h, pb2, vx, vy);
                sto2<NT>(h + 2 * NP, pb2, vz, dxux);
                sto2<NT>(h + 4 * NP, pb2, dyuy, dzuz);
                sto2<NT>(h + 6 * NP, pb2, (dxuy + dyux) / 2.0f, (dxuz + dzux) / 2.0f);
                sto<NT>(h + 8 * NP, pb, (dyuz + dzuy) / 2.0f);
            }
        }
        if (MODE == 2) {
            // kernels (SOURCE/ses3d_grad.f90:53-83, ses3d_grad_Q.f90:115-126); float32
            // increments in the reference's order, accumulated in float64
            const int sl = ldg(E.slot + it);
            if (sl >= 0) {
                float f9[9];
                if (hcomp) {
                    // ses3d_restore with deg < lpd (ses3d_output.f90:242-265): the
                    // stored values interpolated to this node, summed in the order
                    // il (x), im (y), in (z) from 0
                    const int D = Ms.hdeg + 1, NC = Ms.hnc;
                    const float *HLD = tabs + 7 * 25 + 20;
#pragma unroll
                    for (int c = 0; c < 9; ++c) f9[c] = 0.f;
#pragma unroll 1
                    for (int il = 0; il < D; ++il)
#pragma unroll 1
                        for (int im = 0; im < D; ++im)
#pragma unroll 1
                            for (int in = 0; in < D; ++in) {
                                const float w = (HLD[il * 5 + l] * HLD[im * 5 + m]) * HLD[in * 5 + n];
                                const int idx = il + D * (im + D * in);
#pragma unroll
                                for (int c = 0; c < 9; ++c) f9[c] = f9[c] + hv[sub][c * NC + idx] * w;
                            }
                } else {
                    const float *h = E.hist + (size_t)sl * 9 * NP;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const f2v t = ldo2<NT>(h + 2 * q * NP, pb2);
                        f9[2 * q] = t.x;
                        f9[2 * q + 1] = t.y;
                    }
                    f9[8] = ldo<NT>(h + 8 * NP, pb);
                }
                const float vxf = f9[0], vyf = f9[1], vzf = f9[2];
                const float exxf = f9[3], eyyf = f9[4], ezzf = f9[5];
                const float exyf = f9[6], exzf = f9[7], eyzf = f9[8];
                const float exya = 0.5f * (dxuy + dyux), exza = 0.5f * (dxuz + dzux), eyza = 0.5f * (dyuz + dzuy);
                const float tra = (dxux + dyuy) + dzuz, trf = (exxf + eyyf) + ezzf;
                const float ede = ((((dxux * exxf + dyuy * eyyf) + dzuz * ezzf) + (2.0f * exya) * exyf) +
                                   (2.0f * exza) * exzf) + (2.0f * eyza) * eyzf;
                const float rhoinv = ldg(Ms.rhoinv + p);
                const float rho = rhoinv > 0.f ? 1.0f / rhoinv : 0.f;
                const float cs = sqrtf(mu * rhoinv);
                const float cp = sqrtf((lam + 2.0f * mu) * rhoinv);
                const float sdt = E.sdt, Jac = Ms.Jac;
                double *gr = E.grad;
                const float g_cp = (sdt * (((2.0f * rho) * cp) * trf * tra)) / Jac;
                const float g_rho1 = (sdt * ((vxf * vx + vyf * vy) + vzf * vz)) / Jac;
                const float g_rho2 = (sdt * (((cp * cp - (2.0f * cs) * cs) * trf) * tra + ((2.0f * cs) * cs) * ede)) / Jac;
                const float g_csh = (sdt * (((4.0f * rho) * cs) *
                                            (((ede - tra * trf) - (2.0f * exzf) * exza) - (2.0f * eyzf) * eyza))) / Jac;
                const float g_csv = (sdt * (((8.0f * rho) * cs) * (exzf * exza + eyzf * eyza))) / Jac;
                float gqi[3] = {0.f, 0.f, 0.f};
                if (DISS) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float Mxx = Mv[0 * 3 + k], Myy = Mv[1 * 3 + k], Mzz = Mv[2 * 3 + k];
                        const float Mxy = Mv[3 * 3 + k], Mxz = Mv[4 * 3 + k], Myz = Mv[5 * 3 + k];
                        const float md = (((((Mxx * exxf + Myy * eyyf) + Mzz * ezzf) + Mxy * exyf) + Mxz * exzf) +
                                          Myz * eyzf) - (trf * ((Mxx + Myy) + Mzz)) / 3.0f;
                        gqi[k] = ((sdt * md) * Ms.tau_p[k]) / Jac;
                    }
                }
                if (Ms.kf32) {
                    // float32 sums in the reference's statement order (bit for bit its kernels)
                    float *gf = (float *)gr;
                    gf2v *g01 = (gf2v *)(gf + 2 * (size_t)p), *g23 = (gf2v *)(gf + 2 * NP + 2 * (size_t)p);
                    f2v a = *g01, b = *g23;
                    a.y = a.y + g_cp;
                    a.x = (a.x + g_rho1) + g_rho2;
                    b.x = b.x + g_csh;
                    b.y = b.y + g_csv;
                    *g01 = a;
                    *g23 = b;
                    if (DISS) gf[4 * NP + p] = ((gf[4 * NP + p] + gqi[0]) + gqi[1]) + gqi[2];
                } else {
                    gd2v *g01 = (gd2v *)(gr + 2 * (size_t)p), *g23 = (gd2v *)(gr + 2 * NP + 2 * (size_t)p);
                    d2v a = *g01, b = *g23;
                    a.x = (a.x + (double)g_rho1) + (double)g_rho2;
                    a.y = a.y + (double)g_cp;
                    b.x = b.x + (double)g_csh;
                    b.y = b.y + (double)g_csv;
                    *g01 = a;
                    *g23 = b;
                    if (DISS) gr[4 * NP + p] = ((gr[4 * NP + p] + (double)gqi[0]) + (double)gqi[1]) + (double)gqi[2];
                }
            }
        }
    }
    __syncthreads();

    if (MODE == 1 && hcomp && act) {
        // ses3d_store with deg < lpd (ses3d_output.f90:36-130): the degree-4
        // interpolant of each field sampled at the degree-deg knots, each value
        // summed over the 125 nodes in the order ip (x), iq (y), ir (z) from 0
        const int sl = ldg(E.slot + it);
        if (sl >= 0) {
            const int D = Ms.hdeg + 1, NC = Ms.hnc;
            const float *HLC = tabs + 7 * 25;
            float *hd = E.hist + (size_t)sl * Ms.hslot;
#pragma unroll 1
            for (int t2 = nd; t2 < 9 * NC; t2 += TPE) {
                const int c = t2 / NC, idx = t2 - c * NC;
                const int i = idx % D, j = (idx / D) % D, k = idx / (D * D);
                const float *f = hv[sub] + c * TPE;
                float acc = 0.f;
#pragma unroll 1
                for (int ip = 0; ip < 5; ++ip)
#pragma unroll 1
                    for (int iq = 0; iq < 5; ++iq) {
                        const float wxy = HLC[ip * 4 + i] * HLC[iq * 4 + j];
#pragma unroll
                        for (int ir = 0; ir < 5; ++ir)
                            acc = acc + f[ip + 5 * (iq + 5 * ir)] * (wxy * HLC[ir * 4 + k]);
                    }
                hd[((size_t)c * NE + e) * NC + idx] = acc;
            }
        }
    }

    if (act) {
        // weak-form stress divergence (:237-270); outside the layers the second
        // member of each stress pair is read from the first
        const int cyx = pml ? EYX : EXY, czx = pml ? EZX : EXZ, czy = pml ? EZY : EYZ;
        const float zc = ldg(Ms.Z + ek * 5 + n), sa = ldg(Ms.S + ei * 5 + l);
        float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const float wx = ((WDL[l * 5 + q] * zc) * ldg(Ms.S + ei * 5 + q)) * DUM1[m * 5 + n];
            const float wy = (WDL[m * 5 + q] * zc) * DUM2[l * 5 + n];
            const float zq = ldg(Ms.Z + ek * 5 + q);
            const float wz = (((WDL[n * 5 + q] * zq) * zq) * sa) * DUM3[l * 5 + m];
            const int pz_ = l + 5 * (m + 5 * q), py_ = l + 5 * (q + 5 * n), px_ = q + 5 * (m + 5 * n);
            ax = ((ax + ss[sub][czx][pz_] * wz) + ss[sub][cyx][py_] * wy) + ss[sub][EXX][px_] * wx;
            ay = ((ay + ss[sub][czy][pz_] * wz) + ss[sub][EYY][py_] * wy) + ss[sub][EXY][px_] * wx;
            az = ((az + ss[sub][EZZ][pz_] * wz) + ss[sub][EYZ][py_] * wy) + ss[sub][EXZ][px_] * wx;
        }
        // point forces / adjoint sources of this element, after the divergence
        if (E.force_elem) {
            const int2 fe = E.force_elem[e];
            if (fe.y > 0 && nd < 125) subtract_forces(E.forces + fe.x, fe.y, nd, Ms.nt, it, ax, ay, az);
        }
        // tail lanes write the slot's padding, real nodes their fperm position
        const unsigned fq = 4u * (unsigned)(e * EST + (nd < 125 ? fperm(l, m, n) : nd));
        stc2(E.f, 2u * fq, ax, ay);
        stc(E.f + 2 * NP, fq, az);
    }
}

// One section: the step number lives in HBM so a captured step graph replays unchanged.
// NT: the state streams bypass the caches (sections larger than the caches).
template <int DISS, int MODE, int ANISO, int EPB, int NT>
__global__ __launch_bounds__(EPB * TPE)
__attribute__((amdgpu_waves_per_eu(WAVES_MIN(MODE, EPB), 8)))
void k_element(const MeshDev *__restrict__ mesh, const EventDev *__restrict__ evs, const int *itp,
               int blocks_per_event, int remap, const RecDev *__restrict__ recs, int nrec) {
    // the first nrec workgroups sample the receivers (recording steps), the
    // others advance the elements; both only read the velocity
    const int b = (int)blockIdx.x - nrec;
    const RecDev *rec = b < 0 ? recs + blockIdx.x : nullptr;
    const int blk = b < 0 ? 0 : remap ? xcd_remap(b, (int)gridDim.x - nrec) : b;
    element_step<DISS, MODE, ANISO, EPB, NT>(*mesh, evs, *itp, blk, blocks_per_event, rec);
}

// Several sections (meshes of any size, each with its own events and step
// counter) advanced by one launch: the workgroups of section s are
// [eb0, eb0 + bpe * n_events) for the element kernel and
// [nb0, nb0 + ceil(nbpe * 64 * n_events / lanes)) for the node kernel.
struct SectionDev {
    const MeshDev *mesh;
    const EventDev *evs;
    int *itp;
    int n_events;
    int eb0, bpe;     // element kernel: first workgroup, workgroups per event
    int nb0, nbpe;    // node kernel: first workgroup, 4x4x4 node blocks per event
};

// Launch tables are read-only while a kernel runs: reading them through the
// constant address space makes their uniform reads scalar loads (through a
// pointer loaded from memory they would be per-lane flat loads).
typedef __attribute__((address_space(4))) const SectionDev cSectionDev;
typedef __attribute__((address_space(4))) const MeshDev cMeshDev;
typedef __attribute__((address_space(4))) const EventDev cEventDev;

__device__ __forceinline__ int section_of(cSectionDev *secs, int nsec, int b, bool node) {
    int s = 0;
    while (s + 1 < nsec && b >= (node ? secs[s + 1].nb0 : secs[s + 1].eb0)) ++s;
    return s;
}

template <int DISS, int MODE, int ANISO, int EPB, int NT>
__global__ __launch_bounds__(EPB * TPE)
__attribute__((amdgpu_waves_per_eu(WAVES_MIN(MODE, EPB), 8)))
void k_element_group(const SectionDev *__restrict__ secs, int nsec, int remap, const RecDev *__restrict__ recs,
                     int nrec) {
    const int r = (int)blockIdx.x - nrec;
    const int b = r < 0 ? 0 : remap ? xcd_remap(r, (int)gridDim.x - nrec) : r;
    cSectionDev *cs = (cSectionDev *)secs;
    cSectionDev &S = cs[section_of(cs, nsec, b, false)];
    element_step<DISS, MODE, ANISO, EPB, NT>(*(cMeshDev *)S.mesh, (cEventDev *)S.evs, *S.itp, b - S.eb0, S.bpe,
                                             r < 0 ? recs + blockIdx.x : nullptr);
}

// ---------------------------------------------------------------------------
// node kernel: assembly in the reference's order (sharing elements in
// lexicographic (i,j,k) order, :286-310), /MM (:318-320), velocity update
// (:326-328), taper (:366).  One thread per global node; a wavefront owns one
// 4x4x4 node block.
// ---------------------------------------------------------------------------
// One node block `nb` (x fastest, then y, then z) of event E; lane `loc`.
template <typename MT, typename EV>
__device__ __forceinline__ void node_block(MT &Ms, EV &E, const int nb, const int loc) {
    const int NX = Ms.NX, NY = Ms.NY, NZ = Ms.NZ, NBX = Ms.NBX, NBY = Ms.NBY;
    const int byz = nb / NBX;
    const int bx = nb - byz * NBX, bz = byz / NBY, by = byz - bz * NBY;
    const int a = loc & 3, b = (loc >> 2) & 3, c = loc >> 4;
    if ((bx == NX && a) || (by == NY && b) || (bz == NZ && c)) return;
    const size_t NP = (size_t)Ms.NP;
    // sharers per axis in ascending element index: (e0, l0) and, on an
    // element face, (e1, l1); held in scalars so nothing goes to scratch
    int xe0, xl0, nxs;
    if (a) { xe0 = bx; xl0 = a; nxs = 1; }
    else if (bx > 0) { xe0 = bx - 1; xl0 = 4; nxs = bx < NX ? 2 : 1; }
    else { xe0 = bx; xl0 = 0; nxs = 1; }
    int ye0, yl0, nys;
    if (b) { ye0 = by; yl0 = b; nys = 1; }
    else if (by > 0) { ye0 = by - 1; yl0 = 4; nys = by < NY ? 2 : 1; }
    else { ye0 = by; yl0 = 0; nys = 1; }
    int ze0, zl0, nzs;
    if (c) { ze0 = bz; zl0 = c; nzs = 1; }
    else if (bz > 0) { ze0 = bz - 1; zl0 = 4; nzs = bz < NZ ? 2 : 1; }
    else { ze0 = bz; zl0 = 0; nzs = 1; }
    // a second sharer exists only on a face between elements: it is (bx, 0) (likewise for y, z)
    // load the <=8 contributions first (independent loads), then add them in
    // the reference's order: sharers sorted by (i, j, k), i slowest
    float fx[8], fy[8], fz[8];
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                const int sl = ii * 4 + jj * 2 + kk;
                fx[sl] = fy[sl] = fz[sl] = 0.f;
                if (ii < nxs && jj < nys && kk < nzs) {
                    const int ex = ii ? bx : xe0, lx = ii ? 0 : xl0;
                    const int ey = jj ? by : ye0, ly = jj ? 0 : yl0;
                    const int ez = kk ? bz : ze0, lz = kk ? 0 : zl0;
                    const int q = (ex + NX * (ey + NY * ez)) * EST + fperm(lx, ly, lz);
                    // 32-bit byte offsets from the uniform plane bases (the
                    // element-local arrays are below 4 GiB, checked on the host)
                    const f2v t = ldc2(E.f, 8u * (unsigned)q);
                    fx[sl] = t.x;
                    fy[sl] = t.y;
                    fz[sl] = ldc(E.f + 2 * NP, 4u * (unsigned)q);
                }
            }
    float Sx = 0.f, Sy = 0.f, Sz = 0.f;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
                if (ii < nxs && jj < nys && kk < nzs) {
                    const int sl = ii * 4 + jj * 2 + kk;
                    Sx = Sx + fx[sl];
                    Sy = Sy + fy[sl];
                    Sz = Sz + fz[sl];
                }
    // absorbing profile and taper of the global node: the reference copies the
    // element-local value of the last element in loop order (init.f90:690-709)
    float pg = 0.f, tg = 1.f;
    {
        const int exl = bx < NX ? bx : NX - 1, eyl = by < NY ? by : NY - 1, ezl = bz < NZ ? bz : NZ - 1;
        const int lxl = 4 * bx + a - 4 * exl, lyl = 4 * by + b - 4 * eyl, lzl = 4 * bz + c - 4 * ezl;
        const unsigned ix = ldg(Ms.IX + exl * 5 + lxl), iy = ldg(Ms.IY + eyl * 5 + lyl), iz = ldg(Ms.IZ + ezl * 5 + lzl);
        if (ix | iy | iz) {
            const float *lut = Ms.LUT + (((size_t)iz * Ms.KY + iy) * Ms.KX + ix) * 5;
            pg = ldg(lut + 3); tg = ldg(lut + 4);
        }
    }
    const int g = (nb << 6) + loc;
    const float mm = ldnt(Ms.MM + g);
    const float dt = Ms.dt;
    float *v = E.v;
    const float ax = Sx / mm, ay = Sy / mm, az = Sz / mm;
    const f2v vxy = *(cgf2v *)(v + 2 * (size_t)g);
    float vx = vxy.x, vy = vxy.y, vz = ldnt(v + 2 * Ms.NGB + g);
    vx = vx + dt * (ax - pg * vx);
    vy = vy + dt * (ay - pg * vy);
    vz = vz + dt * (az - pg * vz);
    f2v o;
    o.x = tg * vx;
    o.y = tg * vy;
    *(gf2v *)(v + 2 * (size_t)g) = o;
    stnt(v + 2 * Ms.NGB + g, tg * vz);
}

// `gt`: the lane's index among the section's n_events * nodeblocks_per_event * 64
// node lanes (32-bit: n_events * padded nodes stays below 2^31, checked on the host).
template <typename MT, typename ET>
__device__ __forceinline__ void node_update(MT &Ms, ET *__restrict__ evs, const int nodeblocks_per_event,
                                            const int n_events, const int gt) {
    // a wavefront owns one node block: the block and event indices are
    // wavefront-uniform, so their divisions run on the scalar unit
    const int nbt = __builtin_amdgcn_readfirstlane(gt >> 6);
    const int ev = nbt / nodeblocks_per_event;
    if (ev >= n_events) return;
    node_block(Ms, evs[ev], nbt - ev * nodeblocks_per_event, gt & 63);
}

__global__ __launch_bounds__(NODE_MAX_THREADS) void k_node(const MeshDev *__restrict__ mesh,
                                                           const EventDev *__restrict__ evs,
                                                           int nodeblocks_per_event, int n_events, int *itp,
                                                           int remap) {
    // last kernel of a step: advance the step number (no kernel of this launch reads it)
    if (blockIdx.x == 0 && threadIdx.x == 0) *itp = *itp + 1;
    const int b = remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    node_update(*mesh, evs, nodeblocks_per_event, n_events, b * (int)blockDim.x + (int)threadIdx.x);
}

__global__ __launch_bounds__(NODE_MAX_THREADS) void k_node_group(const SectionDev *__restrict__ secs, int nsec,
                                                                 int remap) {
    const int b = remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    cSectionDev *cs = (cSectionDev *)secs;
    cSectionDev &S = cs[section_of(cs, nsec, b, true)];
    if (b == S.nb0 && threadIdx.x == 0) *S.itp = *S.itp + 1;
    node_update(*(cMeshDev *)S.mesh, (cEventDev *)S.evs, S.nbpe, S.n_events,
                (b - S.nb0) * (int)blockDim.x + (int)threadIdx.x);
}

// Count, for each divisor b[d], the float32 numerators a (all 2^32 bit
// patterns except NaN/Inf) for which div_recip(a, b, y) differs from a / b.
__global__ void k_verify_div(const float *b, const float *y, int nd, unsigned long long *bad) {
    const unsigned stride = gridDim.x * blockDim.x;
    for (int d = 0; d < nd; ++d) {
        const float bd = b[d], yd = y[d];
        unsigned long long nbad = 0;
        for (unsigned long long u = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; u < (1ULL << 32); u += stride) {
            const float a = __uint_as_float((unsigned)u);
            if (!__builtin_isfinite(a)) continue;
            const float ref = a / bd;
            const float fast = div_recip(a, bd, yd);
            nbad += (__float_as_uint(ref) != __float_as_uint(fast)) && !(ref == 0.0f && fast == 0.0f);
        }
        if (nbad) atomicAdd(&bad[d], nbad);
    }
}

// One component of an element-major array (point p at src[p * stride]) into
// the reference layout (element fastest, node slowest): dst[e + NE * nd].
// element e's node nd at src[e * es + nd * stride] (es: the element stride)
template <typename T>
__global__ void k_to_ref(T *__restrict__ dst, const T *__restrict__ src, long long NE, int stride, long long es) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NE * 125) return;
    const long long nd = i / NE, e = i - nd * NE;
    dst[i] = src[e * es + nd * stride];
}

// Inverse of k_to_ref: reference layout into one component of an
// element-major array (the 3 padding values of each slot are left as they are).
template <typename T>
__global__ void k_from_ref(T *__restrict__ dst, const T *__restrict__ src, long long NE, int stride, long long es) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NE * 125) return;
    const long long nd = i / NE, e = i - nd * NE;
    dst[e * es + nd * stride] = src[i];
}

// float32 accumulators widened into a float64 reference-layout array
__global__ void k_to_ref_widen(double *__restrict__ dst, const float *__restrict__ src, long long NE, int stride) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NE * 125) return;
    const long long nd = i / NE, e = i - nd * NE;
    dst[i] = (double)src[(e * EST + nd) * stride];
}

// grad_Q_mu <- 2 mu grad_Q_mu / Q at it == nt (SOURCE/ses3d_grad_Q.f90:151-153),
// in the accumulators' type; g in reference layout, mu and Q element-major
__global__ void k_q_normalise(double *__restrict__ g, const float *__restrict__ lm, const float *__restrict__ Q,
                              long long NE, int f32) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NE * 125) return;
    const long long nd = i / NE, e = i - nd * NE, p = e * EST + nd;
    const float mu = lm[2 * p + 1], q = Q[p];
    if (f32) g[i] = (double)(((2.0f * mu) * (float)g[i]) / q);
    else g[i] = ((2.0 * (double)mu) * g[i]) / (double)q;
}

// flag <- 1 if any of a, b, c (n values each) is nonzero
__global__ void k_any_nonzero(const float *__restrict__ a, const float *__restrict__ b, const float *__restrict__ c,
                              long long n, int *flag) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (a[i] != 0.0f || b[i] != 0.0f || c[i] != 0.0f) atomicOr(flag, 1);
}

__global__ void k_set_step(int *itp, int it) {
    if (threadIdx.x == 0) *itp = it;
}

// mass matrix (SOURCE/ses3d_init.f90:254-277): the same ordered gather as k_node
__global__ __launch_bounds__(DEFAULT_NODE_THREADS) void k_mass(MeshDev *mesh, float *MM, const float *w5) {
    const MeshDev &Ms = *mesh;
    const long long gt = (long long)blockIdx.x * DEFAULT_NODE_THREADS + threadIdx.x;
    const long long nb = gt >> 6;
    const int loc = (int)(gt & 63);
    const int NX = Ms.NX, NY = Ms.NY, NZ = Ms.NZ, NBX = Ms.NBX, NBY = Ms.NBY;
    const int bx = (int)(nb % NBX), by = (int)((nb / NBX) % NBY), bz = (int)(nb / ((long long)NBX * NBY));
    const int a = loc & 3, b = (loc >> 2) & 3, c = loc >> 4;
    if (bz >= Ms.NBZ) return;
    if ((bx == NX && a) || (by == NY && b) || (bz == NZ && c)) { MM[gt] = 1.0f; return; }
    int xe[2], xl[2], nxs = 0, ye[2], yl[2], nys = 0, ze[2], zl[2], nzs = 0;
    if (a) { xe[0] = bx; xl[0] = a; nxs = 1; }
    else {
        if (bx > 0) { xe[nxs] = bx - 1; xl[nxs] = 4; ++nxs; }
        if (bx < NX) { xe[nxs] = bx; xl[nxs] = 0; ++nxs; }
    }
    if (b) { ye[0] = by; yl[0] = b; nys = 1; }
    else {
        if (by > 0) { ye[nys] = by - 1; yl[nys] = 4; ++nys; }
        if (by < NY) { ye[nys] = by; yl[nys] = 0; ++nys; }
    }
    if (c) { ze[0] = bz; zl[0] = c; nzs = 1; }
    else {
        if (bz > 0) { ze[nzs] = bz - 1; zl[nzs] = 4; ++nzs; }
        if (bz < NZ) { ze[nzs] = bz; zl[nzs] = 0; ++nzs; }
    }
    float mm = 0.f;
    for (int ii = 0; ii < nxs; ++ii)
        for (int jj = 0; jj < nys; ++jj)
            for (int kk = 0; kk < nzs; ++kk) {
                const int i = xe[ii], j = ye[jj], k = ze[kk], L = xl[ii], Mi = yl[jj], N = zl[kk];
                const long long q = ((long long)i + (long long)NX * (j + (long long)NY * k)) * EST + (L + 5 * (Mi + 5 * N));
                const float zk = Ms.Z[k * 5 + N];
                const float tterm = (((((((w5[L] * w5[Mi]) * w5[N]) * Ms.Jac) * zk) * zk) * Ms.S[i * 5 + L])) / Ms.rhoinv[q];
                mm = mm - tterm;
            }
    MM[gt] = mm;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
};

}  // namespace

struct GraphEntry {
    int mode = 0;
    bool record = false;
    int steps = 0;
    hipGraphExec_t exec = nullptr;
};

// Launch schedule of the element and node kernels.
struct Schedule {
    int epb = DEFAULT_EPB;                    // elements per element-kernel workgroup: 1, 2, 4
    int node_threads = NODE_MAX_THREADS;      // lanes per node-kernel workgroup: 128, 256, 512
    int remap_elem = 0, remap_node = 0;       // XCD-aware workgroup order (measured defaults)
};

constexpr int N_TRIALS = 6;

// The time-step driver shared by one section (ses3d_hip_context) and a group
// of sections advanced together (ses3d_hip_group): stream-ordered steps,
// capture and replay of step graphs, per-kernel timing and the autotuner.
struct StepRunner {
    hipStream_t stream = nullptr;
    Schedule sched;
    int it_host = -1;                 // value the device step counter(s) hold next (-1 unknown)
    int graph_steps = 16;
    int streaming = -1;               // element-kernel state streams: -1 by state size, 0 cached, 1 non-temporal
    std::vector<GraphEntry> graphs;
    bool profiling = false;
    double t_total = 0, t_elem = 0, t_node = 0;
    int n_elem = 0;
    std::vector<hipEvent_t> tev;
    float tune_ms[N_TRIALS][2] = {};  // autotuner: fastest element / node kernel per trial
    bool tuned = false;

    virtual ~StepRunner() {}
    // the kernels of one step for every event, reading the step number from HBM
    // the element kernel of one step; `record` adds the receivers' workgroups
    virtual int launch_element(int mode, bool record) = 0;
    virtual int launch_node() = 0;
    virtual int set_step(int it) = 0;   // store it in every step counter, stream-ordered
    virtual int schedule_changed() {
        drop_graphs();
        return 0;
    }
    // Executable graphs are destroyed only once the stream has drained, so no
    // pending launch of one outlives it.
    void drop_graphs() {
        if (graphs.empty()) return;
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    void release_runner() {
        drop_graphs();
        for (auto &e : tev) (void)hipEventDestroy(e);
        tev.clear();
    }
};

struct EventHost {
    int source_type = 0;
    float xxs = 0, yys = 0, zzs = 0, mom[6] = {0, 0, 0, 0, 0, 0};
    std::vector<float> so;
    bool have_source = false;
    std::vector<float> rec;   // 3*nr
    std::vector<float> ad_loc, ad_stf;   // 3*n, n*nt*3
    // located
    ses3d::SourceLoc sloc;
    std::vector<ses3d::PointLoc> recs, ads;
    // device
    float *v = nullptr, *st = nullptr, *f = nullptr;   // st: NE state blocks (StateLayout)
    float *src_md = nullptr, *so_d = nullptr;
    float *seis = nullptr;     // nr x 3 x nt
    float *rec_w = nullptr;    // nr x 125
    float *force_w = nullptr, *force_stf = nullptr;
    int2 *force_elem = nullptr;   // per element {first, count} into the context's force list
    size_t force_elem_cap = 0;
    bool has_forces = false;
    size_t seis_cap = 0, rec_w_cap = 0, force_w_cap = 0, force_stf_cap = 0;
    double *grad = nullptr;
    float *hist = nullptr;       // sampled forward fields the kernels read/write (active buffer)
    float *hist_own = nullptr;   // full history of every sampled step (allocated on first use)
    int n_slots = 0;
    std::vector<int> slot_h;     // per forward step: index among the sampled steps, or -1
    // restart checkpoints (forward_checkpointed) and the segment buffer replayed into
    std::vector<float *> ckpt;
    float *seg = nullptr;
    int seg_slots = 0;
};

struct ses3d_hip_context : StepRunner {
    ses3d_hip_config cfg;
    int n_events = 0;
    int NX = 0, NY = 0, NZ = 0, NBX = 0, NBY = 0, NBZ = 0;
    long long NE = 0, NP = 0, NGB = 0;   // NP: element-local slots NE*EST (device stride)
    long long NPTS = 0;                  // GLL points NE*125 (reference-layout host arrays)
    int device = 0;
    int generation = 0;               // counts inits (kernel arguments may change)
    ses3d::MeshSetup ms;
    bool model_set = false, initialised = false;
    int hdeg = 4;                        // degree of the stored forward fields (4: unchanged)
    int kernel_bits = 64;                // kernel accumulators: 64 float64, 32 float32 (the reference's)
    // device mesh data
    float *lm = nullptr, *A = nullptr, *B = nullptr, *C = nullptr, *Q = nullptr,
          *rhoinv = nullptr, *MM = nullptr;
    float *Z = nullptr, *S = nullptr, *COT = nullptr, *LUT = nullptr, *w5 = nullptr;
    unsigned short *IX = nullptr, *IY = nullptr, *IZ = nullptr;
    unsigned char *PMLX = nullptr, *PMLY = nullptr, *PMLZ = nullptr;
    std::vector<unsigned char> pmlx, pmly, pmlz;   // host copies
    bool aniso = true;                             // any of A, B, C nonzero
    MeshDev mesh_h{};
    MeshDev *mesh_d = nullptr;
    std::vector<EventHost> evh;
    EventDev *ev_d = nullptr;
    // receivers / forces (all events)
    RecDev *recs_d = nullptr;
    std::vector<RecDev> recs_h;   // the table on the device (groups concatenate them)
    int nrec_total = 0;
    ForceDev *forces_d = nullptr;
    size_t recs_cap = 0, forces_cap = 0;
    std::vector<int> saving_vector;   // 1..nt
    int ckpt_k = 0;
    int ckpt_keep = 1;       // trailing segments whose sampled steps the checkpointed forward keeps
    int ckpt_keep_used = 0;  // min(ckpt_keep, segments) of the last checkpointed forward
    bool ckpt_ready = false;
    bool last_seg_ready = false;   // the segment buffer holds the kept trailing segments' sampled steps
    // step number in HBM
    int *it_d = nullptr;
    unsigned long long div_bad[4] = {0, 0, 0, 0};   // numerators failing the reciprocal check
    std::vector<std::pair<void *, size_t>> allocs;
    int64_t bytes = 0;
    unsigned long long *div_buf = nullptr;        // divisors, reciprocals, failure counts
    float *conv = nullptr;                         // layout-conversion scratch for downloads
    size_t conv_cap = 0;
    int *flag_d = nullptr;                         // device-side anisotropy test of set_model_device
    // [it] -> slot of the sampled-step buffers: the same table for every event
    // of the context (their sampled steps are the same), uploaded once
    int *slot_d = nullptr;
    std::vector<int> slot_dev;                     // the table currently on the device
    std::vector<EventDev> ev_dev;                  // the event table currently on the device

    int set_step_on(hipStream_t st, int it);
    int launch_element(int mode, bool record) override;
    int launch_node() override;
    int set_step(int it) override { return set_step_on(stream, it); }
};

// Sections of any size (each a context with its own events) advanced by one
// element-kernel and one node-kernel launch per step.
struct ses3d_hip_group : StepRunner {
    std::vector<ses3d_hip_context *> ctx;
    std::vector<int> gens;            // member generations the launch table was built for
    SectionDev *secs_d = nullptr;
    int elem_wgs = 0, node_wgs = 0;   // workgroups of one group launch
    RecDev *recs_d = nullptr;         // every member's receivers
    size_t recs_cap = 0;
    int nrec = 0;
    int diss = 0, aniso = 0, af = 0;

    int build_table();
    int schedule_changed() override {
        drop_graphs();
        return build_table();
    }
    int launch_element(int mode, bool record) override;
    int launch_node() override;
    int set_step(int it) override {
        for (auto *c : ctx) {
            int rc = c->set_step_on(stream, it);
            if (rc) return rc;
        }
        return 0;
    }
};

namespace {

template <typename T>
int dalloc(ses3d_hip_context *c, T **p, size_t n) {
    void *q = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return fail(SES3D_HIP_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    c->allocs.push_back({q, bytes});
    c->bytes += (int64_t)bytes;
    *p = (T *)q;
    return 0;
}

void dfree(ses3d_hip_context *c, void *p) {
    for (size_t k = 0; k < c->allocs.size(); ++k)
        if (c->allocs[k].first == p) {
            c->bytes -= (int64_t)c->allocs[k].second;
            c->allocs.erase(c->allocs.begin() + (long)k);
            (void)hipFree(p);
            return;
        }
}

#define TRY(expr)                       \
    do {                                \
        int _rc = (expr);               \
        if (_rc) return _rc;            \
    } while (0)

// (re)allocate *p for n elements unless it already holds that many (inits
// after new receiver or source lists may need larger buffers)
template <typename T>
int dfit(ses3d_hip_context *c, T **p, size_t &cap, size_t n) {
    if (*p && cap >= n) return 0;
    if (*p) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        dfree(c, *p);
        *p = nullptr;
        cap = 0;
    }
    TRY(dalloc(c, p, n));
    cap = n;
    return 0;
}

long long host_gidx(const ses3d_hip_context *c, int gx, int gy, int gz) {
    return ((((long long)(gz >> 2) * c->NBY + (gy >> 2)) * c->NBX + (gx >> 2)) << 6) +
           (((gz & 3) << 4) | ((gy & 3) << 2) | (gx & 3));
}

int check_ev(ses3d_hip_context *c, int ev) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (ev < 0 || ev >= c->n_events) return fail(SES3D_HIP_EINVAL, "event index out of range");
    return 0;
}

int upload(ses3d_hip_context *c, void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int download(ses3d_hip_context *c, void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Component of an element-major device array (first point at `comp`, `stride`
// values between points) to a host array in reference layout, transposed on
// the GPU into a scratch buffer kept by the context.
template <typename T>
int download_ref(ses3d_hip_context *c, T *host, const T *comp, int stride, long long es = -1) {
    const size_t n = (size_t)c->NPTS;
    TRY(dfit(c, &c->conv, c->conv_cap, (n * sizeof(T) + sizeof(float) - 1) / sizeof(float)));
    T *d = (T *)c->conv;
    hipLaunchKernelGGL(k_to_ref<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, d, comp, c->NE, stride,
                       es < 0 ? (long long)EST * stride : es);
    HIP_TRY(hipGetLastError());
    return download(c, host, d, sizeof(T) * n);
}

// Host array in reference layout into one component of an element-major
// device array (first point at `comp`, `stride` values between points).
template <typename T>
int upload_ref(ses3d_hip_context *c, T *comp, const T *host, int stride) {
    const size_t n = (size_t)c->NPTS;
    TRY(dfit(c, &c->conv, c->conv_cap, (n * sizeof(T) + sizeof(float) - 1) / sizeof(float)));
    T *d = (T *)c->conv;
    TRY(upload(c, d, host, sizeof(T) * n));
    hipLaunchKernelGGL(k_from_ref<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, comp, d, c->NE, stride,
                       (long long)EST * stride);
    HIP_TRY(hipGetLastError());
    return 0;
}

int zero(ses3d_hip_context *c, void *p, size_t bytes) {
    // a fill kernel would also do; the context's stream is never captured here
    HIP_TRY(hipMemsetAsync(p, 0, bytes, c->stream));
    return 0;
}

int elem_blocks(const ses3d_hip_context *c, int epb) { return (int)((c->NE + epb - 1) / epb); }
// floats of one event's state blocks
size_t state_floats(const ses3d_hip_context *c) { return (size_t)c->NE * StateLayout(c->cfg.is_diss).esb; }
long long node_blocks(const ses3d_hip_context *c) { return (long long)c->NBX * c->NBY * c->NBZ; }
// Bytes of element-local state one step streams (all events): displacement
// gradient, split stresses, memory variables, element force, coefficients.
// Above the MI355X's 256 MiB last-level cache the element kernel reads and
// writes them non-temporally; below it they stay cached between steps.
constexpr long long CACHED_STATE_BYTES = 256LL << 20;
long long stream_bytes(const ses3d_hip_context *c) {
    const long long per_event = 9 + 9 + (c->cfg.is_diss ? 18 : 0) + 3;
    return 4LL * c->NP * (per_event * c->n_events + 4);
}
bool use_streaming(const StepRunner *r, long long bytes) {
    return r->streaming < 0 ? bytes > CACHED_STATE_BYTES : r->streaming != 0;
}
bool use_streaming(const ses3d_hip_context *c) { return use_streaming(c, stream_bytes(c)); }
bool use_streaming(const ses3d_hip_group *g) {
    long long bytes = 0;
    for (auto *c : g->ctx) bytes += stream_bytes(c);
    return use_streaming(g, bytes);
}
// floats of one sampled step of stored forward fields: 9 fields of NP values,
// or of NE*(hdeg+1)^3 values when stored at a lower degree
long long hist_slot_floats(const ses3d_hip_context *c) {
    const long long D = c->hdeg + 1;
    return c->hdeg < 4 ? 9 * c->NE * D * D * D : 9 * c->NP;
}

// One element-kernel launch: a single section (secs == nullptr) or a group.
struct ElemLaunch {
    hipStream_t stream;
    int epb, remap;
    unsigned wgs;
    const MeshDev *mesh;
    const EventDev *evs;
    const int *itp;
    int bpe;
    const SectionDev *secs;
    int nsec;
    bool nt;    // state streams non-temporal
    const RecDev *recs;   // receivers sampled by the first nrec workgroups (recording steps)
    int nrec;
};

template <int D, int MO, int AN, int EP, int NT>
void launch_elem_nt(const ElemLaunch &L) {
    if (L.secs)
        hipLaunchKernelGGL((k_element_group<D, MO, AN, EP, NT>), dim3(L.wgs + L.nrec), dim3(EP * TPE), 0, L.stream,
                           L.secs, L.nsec, L.remap, L.recs, L.nrec);
    else
        hipLaunchKernelGGL((k_element<D, MO, AN, EP, NT>), dim3(L.wgs + L.nrec), dim3(EP * TPE), 0, L.stream, L.mesh,
                           L.evs, L.itp, L.bpe, L.remap, L.recs, L.nrec);
}

template <int D, int MO, int AN, int EP>
void launch_elem_epb(const ElemLaunch &L) {
    if (L.nt) launch_elem_nt<D, MO, AN, EP, 1>(L);
    else launch_elem_nt<D, MO, AN, EP, 0>(L);
}

template <int D, int MO, int AN>
int launch_elem_variant(const ElemLaunch &L) {
    switch (L.epb) {
        case 1: launch_elem_epb<D, MO, AN, 1>(L); break;
        case 2: launch_elem_epb<D, MO, AN, 2>(L); break;
        case 4: launch_elem_epb<D, MO, AN, 4>(L); break;
        default: return fail(SES3D_HIP_EINVAL, "elements per workgroup must be 1, 2 or 4");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_elem(int diss, int mode, int aniso, const ElemLaunch &L) {
    switch ((diss ? 100 : 0) + mode * 10 + (aniso ? 1 : 0)) {
        case 0: return launch_elem_variant<0, 0, 0>(L);
        case 1: return launch_elem_variant<0, 0, 1>(L);
        case 10: return launch_elem_variant<0, 1, 0>(L);
        case 11: return launch_elem_variant<0, 1, 1>(L);
        case 20: return launch_elem_variant<0, 2, 0>(L);
        case 21: return launch_elem_variant<0, 2, 1>(L);
        case 100: return launch_elem_variant<1, 0, 0>(L);
        case 101: return launch_elem_variant<1, 0, 1>(L);
        case 110: return launch_elem_variant<1, 1, 0>(L);
        case 111: return launch_elem_variant<1, 1, 1>(L);
        case 120: return launch_elem_variant<1, 2, 0>(L);
        case 121: return launch_elem_variant<1, 2, 1>(L);
        default: return fail(SES3D_HIP_EINVAL, "no element kernel for this mode");
    }
}

bool valid_schedule(const Schedule &s) {
    return (s.epb == 1 || s.epb == 2 || s.epb == 4) &&
           (s.node_threads == 128 || s.node_threads == 256 || s.node_threads == 512) &&
           (s.remap_elem == 0 || s.remap_elem == 1) && (s.remap_node == 0 || s.remap_node == 1);
}

// The autotuner's trials: trial k times element schedule (epb, order) k and
// node schedule (lanes, order) k in the same steps; the two kernels are timed
// separately and chosen independently.
Schedule trial_schedule(int k) {
    static const int epb[3] = {1, 2, 4}, lanes[3] = {128, 256, 512};
    Schedule s;
    s.epb = epb[k % 3];
    s.node_threads = lanes[k % 3];
    s.remap_elem = s.remap_node = k / 3;
    return s;
}

// One iteration of the reference time loop for all events, reading the step
// number from HBM (graph-capturable).  With `timed`, HIP events bracket the
// element and node kernels on the runner's stream; they are read back after
// the run, so the stream is never stalled.
int launch_step(StepRunner *r, int mode, bool record, bool timed) {
    hipEvent_t *ev = nullptr;
    if (timed) {
        const size_t need = 4 * (size_t)(r->n_elem + 1);
        while (r->tev.size() < need) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            r->tev.push_back(e);
        }
        ev = &r->tev[4 * (size_t)r->n_elem];
        HIP_TRY(hipEventRecord(ev[0], r->stream));
    }
    TRY(r->launch_element(mode, record));
    if (ev) HIP_TRY(hipEventRecord(ev[1], r->stream));
    if (ev) HIP_TRY(hipEventRecord(ev[2], r->stream));
    TRY(r->launch_node());
    if (ev) {
        HIP_TRY(hipEventRecord(ev[3], r->stream));
        r->n_elem += 1;
    }
    return 0;
}

int step(StepRunner *r, int it, int mode, bool record, bool timed) {
    if (r->it_host != it) TRY(r->set_step(it));
    TRY(launch_step(r, mode, record, timed));
    r->it_host = it + 1;
    return 0;
}

// Steps it_begin.. through graphs of graph_steps captured steps each (one graph
// per (mode, record) pair, captured on first use), the tail eagerly.
int run_steps(StepRunner *r, int it_begin, int n, int mode, bool record) {
    int it = it_begin;
    const int G = r->graph_steps;
    if (G > 0 && !r->profiling && n >= G) {
        GraphEntry *ge = nullptr;
        for (auto &g : r->graphs)
            if (g.mode == mode && g.record == record && g.steps == G) ge = &g;
        if (!ge) {
            GraphEntry g;
            g.mode = mode; g.record = record; g.steps = G;
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeRelaxed));
            int rc = 0;
            for (int k = 0; k < G && !rc; ++k) rc = launch_step(r, mode, record, false);
            hipError_t e = hipStreamEndCapture(r->stream, &graph);
            // the captured graph is released on every path below
            struct GraphGuard {
                hipGraph_t g;
                ~GraphGuard() {
                    if (g) (void)hipGraphDestroy(g);
                }
            } guard{e == hipSuccess ? graph : nullptr};
            if (rc) return rc;
            if (e != hipSuccess) return fail(SES3D_HIP_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
            HIP_TRY(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
            r->graphs.push_back(g);
            ge = &r->graphs.back();
        }
        if (r->it_host != it) TRY(r->set_step(it));
        while (it + G <= it_begin + n) {
            HIP_TRY(hipGraphLaunch(ge->exec, r->stream));
            it += G;
        }
        r->it_host = it;
    }
    for (; it < it_begin + n; ++it) TRY(step(r, it, mode, record, r->profiling));
    return 0;
}

int collect_timing(StepRunner *r) {
    if (!r->profiling) return 0;
    HIP_TRY(hipStreamSynchronize(r->stream));
    for (int k = 0; k < r->n_elem; ++k) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, r->tev[4 * k], r->tev[4 * k + 1]));
        HIP_TRY(hipEventElapsedTime(&b, r->tev[4 * k + 2], r->tev[4 * k + 3]));
        r->t_elem += a;
        r->t_node += b;
    }
    return 0;
}

// Steps it_begin..it_begin+n-1 as run_steps would take them, the first
// N_TRIALS * (n / N_TRIALS) of them eagerly under the trial schedules (every
// schedule gives bit-identical results), the rest under the fastest element
// and node schedules found (the minimum kernel time over each trial's steps).
int autotune_steps(StepRunner *r, int it_begin, int n, int mode, bool record) {
    const int per = n / N_TRIALS;
    if (per < 1) return fail(SES3D_HIP_EINVAL, "autotune needs at least 6 steps");
    const Schedule keep = r->sched;
    int it = it_begin;
    r->n_elem = 0;
    for (int k = 0; k < N_TRIALS; ++k) {
        r->sched = trial_schedule(k);
        TRY(r->schedule_changed());
        for (int j = 0; j < per; ++j, ++it) TRY(step(r, it, mode, record, true));
    }
    HIP_TRY(hipStreamSynchronize(r->stream));
    int be = 0, bn = 0;
    for (int k = 0; k < N_TRIALS; ++k) {
        float te = 1e30f, tn = 1e30f;
        for (int j = 0; j < per; ++j) {
            const size_t q = 4 * (size_t)(k * per + j);
            float a = 0, b = 0;
            HIP_TRY(hipEventElapsedTime(&a, r->tev[q], r->tev[q + 1]));
            HIP_TRY(hipEventElapsedTime(&b, r->tev[q + 2], r->tev[q + 3]));
            te = std::min(te, a);
            tn = std::min(tn, b);
        }
        r->tune_ms[k][0] = te;
        r->tune_ms[k][1] = tn;
        if (te < r->tune_ms[be][0]) be = k;
        if (tn < r->tune_ms[bn][1]) bn = k;
    }
    Schedule s = keep;
    const Schedule se = trial_schedule(be), sn = trial_schedule(bn);
    s.epb = se.epb; s.remap_elem = se.remap_elem;
    s.node_threads = sn.node_threads; s.remap_node = sn.remap_node;
    r->sched = s;
    TRY(r->schedule_changed());
    r->tuned = true;
    r->n_elem = 0;
    return run_steps(r, it, it_begin + n - it, mode, record);
}

}  // namespace

int ses3d_hip_context::set_step_on(hipStream_t st, int it) {
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, st, it_d, it);
    HIP_TRY(hipGetLastError());
    return 0;
}

int ses3d_hip_context::launch_element(int mode, bool record) {
    const int bpe = elem_blocks(this, sched.epb);
    ElemLaunch L{stream, sched.epb, sched.remap_elem, (unsigned)(bpe * n_events), mesh_d, ev_d, it_d, bpe, nullptr, 0,
                 use_streaming(this), recs_d, record ? nrec_total : 0};
    return launch_elem(cfg.is_diss, mode, aniso ? 1 : 0, L);
}

int ses3d_hip_context::launch_node() {
    const long long threads = node_blocks(this) * 64 * n_events;
    const int lanes = sched.node_threads;
    hipLaunchKernelGGL(k_node, dim3((unsigned)((threads + lanes - 1) / lanes)), dim3(lanes), 0, stream, mesh_d, ev_d,
                       (int)node_blocks(this), n_events, it_d, sched.remap_node);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Launch table of a group for its current schedule: sections in member order,
// each a contiguous range of workgroups.
int ses3d_hip_group::build_table() {
    std::vector<SectionDev> h(ctx.size());
    long long eb = 0, nb = 0;
    for (size_t k = 0; k < ctx.size(); ++k) {
        ses3d_hip_context *c = ctx[k];
        SectionDev &S = h[k];
        S.mesh = c->mesh_d; S.evs = c->ev_d; S.itp = c->it_d; S.n_events = c->n_events;
        S.bpe = elem_blocks(c, sched.epb);
        S.eb0 = (int)eb;
        eb += (long long)S.bpe * c->n_events;
        S.nbpe = (int)node_blocks(c);
        S.nb0 = (int)nb;
        nb += (node_blocks(c) * 64 * c->n_events + sched.node_threads - 1) / sched.node_threads;
        if (eb >= (1LL << 31) || nb >= (1LL << 31)) return fail(SES3D_HIP_EINVAL, "group too large for one launch");
    }
    elem_wgs = (int)eb;
    node_wgs = (int)nb;
    // every member's receivers, sampled by the first workgroups of the element launch
    std::vector<RecDev> r;
    for (auto *c : ctx) r.insert(r.end(), c->recs_h.begin(), c->recs_h.end());
    if ((eb + (long long)r.size()) * sched.epb * TPE >= (1LL << 32) || nb * sched.node_threads >= (1LL << 32))
        return fail(SES3D_HIP_EINVAL, "group too large for one launch");
    if (r.size() > recs_cap) {
        HIP_TRY(hipStreamSynchronize(stream));   // no launch of the group still reads the old table
        if (recs_d) HIP_TRY(hipFree(recs_d));
        recs_d = nullptr;
        recs_cap = 0;
        if (hipMalloc((void **)&recs_d, sizeof(RecDev) * r.size()) != hipSuccess)
            return fail(SES3D_HIP_ENOMEM, "hipMalloc of the group receiver table failed");
        recs_cap = r.size();
    }
    nrec = (int)r.size();
    if (!r.empty())
        HIP_TRY(hipMemcpyAsync(recs_d, r.data(), sizeof(RecDev) * r.size(), hipMemcpyHostToDevice, stream));
    gens.clear();
    for (auto *c : ctx) gens.push_back(c->generation);
    HIP_TRY(hipMemcpyAsync(secs_d, h.data(), sizeof(SectionDev) * h.size(), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
}

int ses3d_hip_group::launch_element(int mode, bool record) {
    ElemLaunch L{stream, sched.epb, sched.remap_elem, (unsigned)elem_wgs, nullptr, nullptr, nullptr, 0, secs_d,
                 (int)ctx.size(), use_streaming(this), recs_d, record ? nrec : 0};
    return launch_elem(diss, mode, aniso, L);
}

int ses3d_hip_group::launch_node() {
    hipLaunchKernelGGL(k_node_group, dim3((unsigned)node_wgs), dim3(sched.node_threads), 0, stream, secs_d,
                       (int)ctx.size(), sched.remap_node);
    HIP_TRY(hipGetLastError());
    return 0;
}

namespace {

int refresh_events(ses3d_hip_context *c);

// A checkpointed forward pass belongs to the model, events and stored-field
// layout it ran with and to the step state it left: anything that changes
// those (a new model, init, another run of the forward context) retires it,
// and the adjoint then refuses to start until forward_checkpointed runs again.
void invalidate_checkpoints(ses3d_hip_context *c) {
    c->ckpt_ready = false;
    c->last_seg_ready = false;
}

// run_steps or autotune_steps, with whole-run and per-kernel timing when profiling
int timed_run(StepRunner *r, int it_begin, int n, int mode, bool record, bool tune) {
    r->t_elem = r->t_node = r->t_total = 0;
    r->n_elem = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (r->profiling) {
        HIP_TRY(hipEventCreate(&t0));
        HIP_TRY(hipEventCreate(&t1));
        HIP_TRY(hipEventRecord(t0, r->stream));
    }
    int rc = tune ? autotune_steps(r, it_begin, n, mode, record) : run_steps(r, it_begin, n, mode, record);
    if (!rc && r->profiling) {
        rc = hipEventRecord(t1, r->stream) == hipSuccess ? collect_timing(r)
                                                         : fail(SES3D_HIP_EDEVICE, "hipEventRecord failed");
        float ms = 0;
        if (!rc && hipEventElapsedTime(&ms, t0, t1) == hipSuccess) r->t_total = ms;
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    return rc;
}

int set_schedule(StepRunner *r, const int32_t *v) {
    Schedule s;
    s.epb = v[0]; s.node_threads = v[1]; s.remap_elem = v[2]; s.remap_node = v[3];
    if (!valid_schedule(s))
        return fail(SES3D_HIP_EINVAL, "schedule: epb in {1,2,4}, node lanes in {128,256,512}, orders in {0,1}");
    HIP_TRY(hipStreamSynchronize(r->stream));
    r->sched = s;
    return r->schedule_changed();
}

int set_streaming(StepRunner *r, int32_t mode) {
    if (mode < -1 || mode > 1)
        return fail(SES3D_HIP_EINVAL, "streaming: -1 (by state size), 0 (cached) or 1 (non-temporal)");
    HIP_TRY(hipStreamSynchronize(r->stream));
    r->streaming = mode;
    r->drop_graphs();
    return 0;
}

void get_schedule(const StepRunner *r, int32_t *v, float *trial_ms) {
    v[0] = r->sched.epb; v[1] = r->sched.node_threads; v[2] = r->sched.remap_elem; v[3] = r->sched.remap_node;
    if (trial_ms)
        for (int k = 0; k < N_TRIALS; ++k) {
            trial_ms[2 * k] = r->tuned ? r->tune_ms[k][0] : -1.f;
            trial_ms[2 * k + 1] = r->tuned ? r->tune_ms[k][1] : -1.f;
        }
}

int build_slots(ses3d_hip_context *c, EventHost &E, bool allocate_history) {
    const int nt = c->cfg.nt, sa = c->cfg.samp_ad > 0 ? c->cfg.samp_ad : 1;
    E.slot_h.assign(nt + 1, -1);
    int n = 0;
    for (int it = 1; it <= nt; ++it)
        if ((it % sa) == 0 || it == 1 || it == nt) E.slot_h[it] = n++;
    E.n_slots = n;
    if (!c->slot_d) TRY(dalloc(c, &c->slot_d, nt + 1));
    if (allocate_history && !E.hist_own) TRY(dalloc(c, &E.hist_own, (size_t)n * hist_slot_floats(c)));
    if (allocate_history) E.hist = E.hist_own;
    return 0;
}

// Device slot table of one event: fwd tables map it -> slot of step it,
// adjoint tables map it -> slot of the paired forward step nt+1-it.
int upload_slots(ses3d_hip_context *c, const std::vector<int> &tab) {
    if (tab == c->slot_dev) return 0;
    c->slot_dev = tab;
    HIP_TRY(hipMemcpyAsync(c->slot_d, c->slot_dev.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// full history of all sampled steps for the stored-history (disk-equivalent) path
int ensure_full_history(ses3d_hip_context *c) {
    const int nt = c->cfg.nt, af = c->cfg.adjoint_flag;
    bool changed = false;
    for (auto &E : c->evh) {
        if (!E.hist_own) {
            TRY(build_slots(c, E, true));
            changed = true;
        }
        if (E.hist != E.hist_own) { E.hist = E.hist_own; changed = true; }
        std::vector<int> sl(nt + 1, -1);
        for (int it = 1; it <= nt; ++it) sl[it] = af == 1 ? E.slot_h[it] : E.slot_h[nt + 1 - it];
        TRY(upload_slots(c, sl));
    }
    if (changed) TRY(refresh_events(c));
    return 0;
}

int refresh_events(ses3d_hip_context *c) {
    std::vector<EventDev> h(c->n_events);
    for (int ev = 0; ev < c->n_events; ++ev) {
        EventHost &E = c->evh[ev];
        EventDev &D = h[ev];
        D.v = E.v; D.st = E.st; D.f = E.f;
        D.src_elem = -1;
        if (E.have_source && E.source_type == 10 && E.sloc.is_source &&
            (c->cfg.adjoint_flag == 0 || c->cfg.adjoint_flag == 1))
            D.src_elem = (int)(E.sloc.isx + (long long)c->NX * (E.sloc.isy + (long long)c->NY * E.sloc.isz));
        D.src_md = E.src_md;
        D.so = E.so_d;
        D.hist = E.hist;
        D.slot = c->slot_d;
        D.grad = E.grad;
        D.sdt = (float)c->cfg.samp_ad * c->cfg.dt;
        D.force_elem = E.has_forces ? E.force_elem : nullptr;
        D.forces = E.has_forces ? c->forces_d : nullptr;
    }
    if (c->ev_dev.size() == h.size() && std::memcmp(c->ev_dev.data(), h.data(), sizeof(EventDev) * h.size()) == 0)
        return 0;   // the table on the device is this one
    c->ev_dev = h;
    return upload(c, c->ev_d, h.data(), sizeof(EventDev) * c->n_events);
}

}  // namespace

// ============================================================================
// kernel projection onto a block grid (TOOLS/SOURCE/project_kernel.f90)
// ============================================================================
namespace {

// Integral of the GLL Lagrange polynomial idx (degree 4) over [lo, hi] of the
// reference interval, float32, as TOOLS/SOURCE/project_kernel.f90:5-163 forms
// it: coefficients of prod (x - knot_i) by the recurrence, powers by binary
// exponentiation, terms summed in ascending degree, divided by the product of
// knot differences.
float int_lag(int idx, const float *knots, float lo, float hi) {
    float z0[5];
    int j = 1;
    for (int i = 0; i <= 4; ++i)
        if (i != idx) z0[j++] = -knots[i];
    float a[5] = {0, 0, 0, 0, 0};
    a[2] = 1.0f;
    a[1] = z0[1] + z0[2];
    a[0] = z0[1] * z0[2];
    for (int i = 3; i <= 4; ++i) {
        for (int k = i - 2; k >= 1; --k) a[k] = a[k] * z0[i] + a[k - 1];
        a[i] = 1.0f;
        a[i - 1] = 0.0f;
        a[0] = 1.0f;
        for (int k = 1; k <= i; ++k) {
            a[i - 1] = a[i - 1] + z0[k];
            a[0] = a[0] * z0[k];
        }
    }
    float norm = 1.0f;
    for (int i = 0; i <= 4; ++i)
        if (i != idx) norm = norm * (knots[idx] - knots[i]);
    auto pw = [](float x, int n) {   // binary exponentiation, n >= 1
        float r = 1.0f;
        while (true) {
            if (n & 1) r = r * x;
            n /= 2;
            if (n == 0) break;
            x = x * x;
        }
        return r;
    };
    float v = 0.0f;
    for (int i = 0; i <= 4; ++i) v = v + ((pw(hi, i + 1) - pw(lo, i + 1)) * a[i]) / (float)(i + 1);
    return v / norm;
}

// index (element) of the GLL coordinate nearest to b, first one in the
// column-major order of c(0:n, 0:4) (Fortran minloc)
int nearest_element(const std::vector<float> &c, int n, float b) {
    int best = 0;
    float bd = INFINITY;
    for (int q = 0; q <= 4; ++q)
        for (int i = 0; i <= n; ++i) {
            const float d = fabsf(c[i * 5 + q] - b);
            if (d < bd) { bd = d; best = i; }
        }
    return best;
}

// Per block of one axis: the elements it integrates over and, per element,
// the 5 integrals of the GLL polynomials between the block's limits in local
// coordinates (project_kernel.f90:575-690).  Lateral axes count coordinates
// upwards from cmin; the radial axis (down=true) indexes elements downwards
// from zmax.
struct AxisTerms {
    std::vector<int> start, elem;
    std::vector<float> w;
};

AxisTerms axis_terms(const std::vector<float> &c, int n, float cmin, float cmax, float d, const float *knots,
                     const std::vector<float> &edge, bool down) {
    AxisTerms T;
    const int nb = (int)edge.size() - 1;
    T.start.push_back(0);
    for (int b = 0; b < nb; ++b) {
        const float lo = edge[b], hi = edge[b + 1];
        if (hi >= cmin && lo <= cmax) {
            int i0, i1;
            if (!down) {
                if (lo > cmin) {
                    i0 = nearest_element(c, n, lo);
                    if (c[i0 * 5 + 4] == lo) i0 += 1;
                } else {
                    i0 = 0;
                }
                if (hi < cmax) {
                    i1 = nearest_element(c, n, hi);
                    if (c[i1 * 5 + 0] == hi) i1 -= 1;
                } else {
                    i1 = n;
                }
            } else {
                i1 = lo > cmin ? nearest_element(c, n, lo) : n;
                i0 = hi < cmax ? nearest_element(c, n, hi) : 0;
            }
            for (int i = i0; i <= i1; ++i) {
                const float c0 = c[i * 5 + 0], c4 = c[i * 5 + 4];
                float la, lb;
                if (!down) {
                    la = lo > c0 ? ((2.0f * lo - c4) - c0) / d : -1.0f;
                    lb = hi < c4 ? ((2.0f * hi - c4) - c0) / d : 1.0f;
                } else {
                    lb = lo > c4 ? ((c0 + c4) - 2.0f * lo) / d : 1.0f;
                    la = hi < c0 ? ((c0 + c4) - 2.0f * hi) / d : -1.0f;
                }
                T.elem.push_back(i);
                for (int q = 0; q < 5; ++q) T.w.push_back(int_lag(q, knots, la, lb));
            }
        }
        T.start.push_back((int)T.elem.size());
    }
    return T;
}

struct ProjAxis {
    const int *start, *elem;
    const float *w;
};

// One thread per block: sum over the block's elements (x, y, z order) and
// their 125 nodes (l, m, n order) of u * intx * inty * intz / vol, float32, in
// the order of project_kernel.f90:694-711.  u: reference layout.
__global__ void k_project(const float *__restrict__ u, int NX, int NY, long long NE, ProjAxis X, ProjAxis Y,
                          ProjAxis Z, int nbx, int nby, int nbz, float vol, float *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbx * nby * nbz) return;
    const int bz = t % nbz, by = (t / nbz) % nby, bx = t / (nbz * nby);
    float acc = 0.0f;
    for (int ix = X.start[bx]; ix < X.start[bx + 1]; ++ix)
        for (int iy = Y.start[by]; iy < Y.start[by + 1]; ++iy)
            for (int iz = Z.start[bz]; iz < Z.start[bz + 1]; ++iz) {
                const long long e = X.elem[ix] + (long long)NX * (Y.elem[iy] + (long long)NY * Z.elem[iz]);
                for (int l = 0; l < 5; ++l) {
                    const float wx = X.w[ix * 5 + l];
                    for (int m = 0; m < 5; ++m) {
                        const float wy = Y.w[iy * 5 + m];
                        for (int n = 0; n < 5; ++n) {
                            const float wz = Z.w[iz * 5 + n];
                            acc = acc + (((u[e + NE * (l + 5 * (m + 5 * n))] * wx) * wy) * wz) / vol;
                        }
                    }
                }
            }
    out[t] = acc;
}

// Sampled-step buffer of every event for a checkpointed forward: the sampled
// steps of the longest segment of k steps, or of the `keep` trailing segments
// the forward pass stores directly, whichever is more (segment_table maps a
// range of steps into it).
int ensure_segment_buffer(ses3d_hip_context *f, int k, int keep) {
    const int nt = f->cfg.nt, nseg = (nt + k - 1) / k;
    int max_slots = 0;
    const EventHost &E0 = f->evh[0];
    for (int sgi = 0; sgi < nseg; ++sgi) {
        int n = 0;
        for (int it = sgi * k + 1; it <= std::min((sgi + 1) * k, nt); ++it) n += E0.slot_h[it] >= 0;
        max_slots = std::max(max_slots, n);
    }
    int tail = 0;
    for (int it = (nseg - keep) * k + 1; it <= nt; ++it) tail += E0.slot_h[it] >= 0;
    max_slots = std::max(max_slots, tail);
    for (auto &E : f->evh)
        if (E.seg_slots < max_slots) {
            if (E.seg) {
                HIP_TRY(hipStreamSynchronize(f->stream));
                dfree(f, E.seg);
            }
            TRY(dalloc(f, &E.seg, (size_t)max_slots * hist_slot_floats(f)));
            E.seg_slots = max_slots;
        }
    return 0;
}

std::vector<int> segment_table(const EventHost &E, int nt, int s0, int s1) {
    std::vector<int> tab(nt + 1, -1);
    int n = 0;
    for (int it = s0; it <= s1; ++it)
        if (E.slot_h[it] >= 0) tab[it] = n++;
    return tab;
}

// Restart states (one per segment of k steps) of every event and the
// sampled-step buffer for `keep` kept segments; buffers already large enough
// are kept, so a repeated checkpointed forward allocates nothing.
int ensure_checkpoints(ses3d_hip_context *c, int k, int keep) {
    const int nseg = (c->cfg.nt + k - 1) / k;
    const size_t NP = (size_t)c->NP, NGB = (size_t)c->NGB;
    const size_t ck = 3 * NGB + state_floats(c);
    for (auto &E : c->evh)
        if ((int)E.ckpt.size() != nseg) {
            HIP_TRY(hipStreamSynchronize(c->stream));
            for (float *q : E.ckpt) dfree(c, q);
            E.ckpt.assign(nseg, nullptr);
            for (int sgi = 0; sgi < nseg; ++sgi) TRY(dalloc(c, &E.ckpt[sgi], ck));
        }
    return ensure_segment_buffer(c, k, keep);
}

}  // namespace

// ============================================================================
// scratch of the no-allocation projection entry point
// ============================================================================
// Device memory of ses3d_hip_project_noalloc: the staged field, the block
// tables and the staged output of one projection.  Each GPU the library runs
// on holds its own copy until the process ends; calls on one GPU share it.
#define SES3D_NOALLOC_SCRATCH_BYTES (64LL << 20)
static __device__ __align__(256) unsigned char scratch[SES3D_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;   // device address of `scratch`, per device

// Looked up once per device, with that device current (ses3d_hip_create);
// host code uses only these addresses, never the array's name.
int scratch_lookup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return fail(SES3D_HIP_EDEVICE, "invalid device index");
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (g_scratch_ptr[(size_t)device]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SES3D_HIP_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[(size_t)device] = p;
    return 0;
}

void *scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && (size_t)device < g_scratch_ptr.size() ? g_scratch_ptr[(size_t)device] : nullptr;
}

// Outcome of the exhaustive reciprocal-quotient check (k_verify_div) per
// device and divisor bit pattern: a pure function of the divisor, its
// reciprocal RN(1/b) and this build, so a process runs it once per divisor
// and device instead of at every init.
std::mutex g_div_mu;
std::map<std::pair<int, uint32_t>, unsigned long long> g_div_bad;

bool div_cached(int device, float b, unsigned long long *bad) {
    uint32_t u;
    std::memcpy(&u, &b, 4);
    std::lock_guard<std::mutex> lock(g_div_mu);
    const auto it = g_div_bad.find({device, u});
    if (it == g_div_bad.end()) return false;
    *bad = it->second;
    return true;
}

void div_store(int device, float b, unsigned long long bad) {
    uint32_t u;
    std::memcpy(&u, &b, 4);
    std::lock_guard<std::mutex> lock(g_div_mu);
    g_div_bad[{device, u}] = bad;
}

// A projection's block tables and its device buffer's extent: the field
// (reference layout), the output, then per axis the block starts, element
// indices and integrals.
struct ProjPlan {
    AxisTerms T[3];
    long long nb = 0;
    float vol = 0;
    size_t bytes = 0;
};

int plan_projection(const ses3d_hip_context *c, int32_t nbx, const float *bx_deg, int32_t nby, const float *by_deg,
                    int32_t nbz, const float *bz_km, ProjPlan &P) {
    if (nbx < 2 || nby < 2 || nbz < 2) return fail(SES3D_HIP_EINVAL, "a block grid needs at least 2 edges per axis");
    const ses3d::MeshSetup &ms = c->ms;
    // edges as project_kernel.f90:381-388 converts them (pi of SOURCE/ses3d_modules.f90:25)
    const float pi = 3.1415926535898f;
    std::vector<float> ex(nbx), ey(nby), ez(nbz);
    for (int i = 0; i < nbx; ++i) ex[i] = (bx_deg[i] * pi) / 180.0f;
    for (int i = 0; i < nby; ++i) ey[i] = (by_deg[i] * pi) / 180.0f;
    for (int i = 0; i < nbz; ++i) ez[i] = bz_km[i] * 1000.0f;
    P.vol = ((bx_deg[1] - bx_deg[0]) * (by_deg[1] - by_deg[0])) * (bz_km[1] - bz_km[0]);
    P.T[0] = axis_terms(ms.x, ms.nx, ms.xmin, ms.xmax, ms.dx, ms.knots, ex, false);
    P.T[1] = axis_terms(ms.y, ms.ny, ms.ymin, ms.ymax, ms.dy, ms.knots, ey, false);
    P.T[2] = axis_terms(ms.z, ms.nz, ms.zmin, ms.zmax, ms.dz, ms.knots, ez, true);
    P.nb = (long long)(nbx - 1) * (nby - 1) * (nbz - 1);
    if (P.nb >= (1LL << 31)) return fail(SES3D_HIP_EINVAL, "block grid too large");
    P.bytes = sizeof(float) * ((size_t)c->NPTS + (size_t)P.nb);
    for (const auto &t : P.T) P.bytes += sizeof(int) * (t.start.size() + t.elem.size() + 1) + sizeof(float) * (t.w.size() + 1);
    return 0;
}

// Stage the field and the tables in `buf` (P.bytes of device memory), run
// k_project on the context's stream and copy the block values to `out`.
int run_projection(ses3d_hip_context *c, const ProjPlan &P, const float *field, int32_t nbx, int32_t nby,
                   int32_t nbz, float *out, char *buf) {
    int rc = 0;
    float *u_d = (float *)buf;
    float *out_d = u_d + c->NPTS;
    int *ints = (int *)(out_d + P.nb);
    ProjAxis PA[3];
    for (int a = 0; a < 3; ++a) {
        const AxisTerms &t = P.T[a];
        int *st = ints;
        int *el = st + t.start.size();
        float *w = (float *)(el + t.elem.size() + 1);
        ints = (int *)(w + t.w.size() + 1);
        PA[a] = {st, el, w};
        if (!rc && hipMemcpyAsync(st, t.start.data(), sizeof(int) * t.start.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = 1;
        if (!rc && !t.elem.empty() && hipMemcpyAsync(el, t.elem.data(), sizeof(int) * t.elem.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = 1;
        if (!rc && !t.w.empty() && hipMemcpyAsync(w, t.w.data(), sizeof(float) * t.w.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = 1;
    }
    if (!rc && hipMemcpyAsync(u_d, field, sizeof(float) * c->NPTS, hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = 1;
    if (!rc) {
        hipLaunchKernelGGL(k_project, dim3((unsigned)((P.nb + 127) / 128)), dim3(128), 0, c->stream, u_d, c->NX, c->NY,
                           c->NE, PA[0], PA[1], PA[2], nbx - 1, nby - 1, nbz - 1, P.vol, out_d);
        if (hipGetLastError() != hipSuccess) rc = 1;
    }
    if (!rc && hipMemcpyAsync(out, out_d, sizeof(float) * P.nb, hipMemcpyDeviceToHost, c->stream) != hipSuccess) rc = 1;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = 1;
    (void)hipStreamSynchronize(c->stream);
    if (rc) return fail(SES3D_HIP_EDEVICE, "kernel projection failed on the device");
    return 0;
}

}  // namespace

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

const char *ses3d_hip_last_error(void) { return g_err.c_str(); }

int ses3d_hip_create(const ses3d_hip_config *cfg, int32_t n_events, ses3d_hip_context **out) {
    if (!cfg || !out) return fail(SES3D_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->lpd != 4) return fail(SES3D_HIP_EINVAL, "only lpd=4 is supported");
    if (cfg->nx < 0 || cfg->ny < 0 || cfg->nz < 0) return fail(SES3D_HIP_EINVAL, "negative element count");
    if (cfg->nt < 1) return fail(SES3D_HIP_EINVAL, "nt must be >= 1");
    if (n_events < 1) return fail(SES3D_HIP_EINVAL, "n_events must be >= 1");
    if (cfg->pml < 0 || cfg->pml > cfg->nx + 1 || cfg->pml > cfg->ny + 1 || cfg->pml > cfg->nz + 1)
        return fail(SES3D_HIP_EINVAL, "pml width exceeds the mesh");
    if (cfg->adjoint_flag < 0 || cfg->adjoint_flag > 2) return fail(SES3D_HIP_EINVAL, "adjoint_flag must be 0, 1 or 2");
    {
        const long long ne = (long long)(cfg->nx + 1) * (cfg->ny + 1) * (cfg->nz + 1);
        // element-local arrays are addressed with 32-bit byte offsets
        // (pair planes hold 2 floats per point)
        if ((ne * EST + CSPAD) * 8 >= (1LL << 32))
            return fail(SES3D_HIP_EINVAL, "more than 4.1 million elements in one section");
        const long long ngb = (long long)(cfg->nx + 2) * (cfg->ny + 2) * (cfg->nz + 2) * 64;
        if (ngb * n_events >= (1LL << 31))
            return fail(SES3D_HIP_EINVAL, "events x global nodes exceed 2^31: split the events over contexts");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(SES3D_HIP_EDEVICE, "no HIP device available");
    auto *c = new ses3d_hip_context();
    c->cfg = *cfg;
    c->n_events = n_events;
    c->NX = cfg->nx + 1; c->NY = cfg->ny + 1; c->NZ = cfg->nz + 1;
    c->NBX = c->NX + 1; c->NBY = c->NY + 1; c->NBZ = c->NZ + 1;
    c->NE = (long long)c->NX * c->NY * c->NZ;
    c->NP = c->NE * EST + CSPAD;
    c->NPTS = c->NE * 125;
    c->NGB = (long long)c->NBX * c->NBY * c->NBZ * 64;
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SES3D_HIP_EDEVICE, "hipStreamCreate failed");
    }
    {
        const int rc = scratch_lookup(c->device);   // the device is current here
        if (rc) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            return rc;
        }
    }
    c->evh.resize(n_events);
    c->ms.build(cfg->nx, cfg->ny, cfg->nz, cfg->pml, cfg->xmin, cfg->xmax, cfg->ymin, cfg->ymax, cfg->zmin,
                cfg->zmax, cfg->zmax_global, cfg->dt);
    int rc = 0;
    const long long NP = c->NP, NGB = c->NGB;
    rc = rc ? rc : dalloc(c, &c->lm, 2 * NP);
    rc = rc ? rc : dalloc(c, &c->A, NP);
    rc = rc ? rc : dalloc(c, &c->B, NP);
    rc = rc ? rc : dalloc(c, &c->C, NP);
    rc = rc ? rc : dalloc(c, &c->rhoinv, NP);
    if (cfg->is_diss) rc = rc ? rc : dalloc(c, &c->Q, NP);
    rc = rc ? rc : dalloc(c, &c->MM, NGB);
    rc = rc ? rc : dalloc(c, &c->mesh_d, 1);
    rc = rc ? rc : dalloc(c, &c->it_d, 1);
    rc = rc ? rc : dalloc(c, &c->ev_d, n_events);
    for (int ev = 0; ev < n_events && !rc; ++ev) {
        EventHost &E = c->evh[ev];
        rc = rc ? rc : dalloc(c, &E.v, 3 * NGB);
        rc = rc ? rc : dalloc(c, &E.st, state_floats(c));
        rc = rc ? rc : dalloc(c, &E.f, 3 * NP);
        rc = rc ? rc : dalloc(c, &E.src_md, 9 * 125);
        rc = rc ? rc : dalloc(c, &E.so_d, cfg->nt);
        if (cfg->adjoint_flag == 2) rc = rc ? rc : dalloc(c, &E.grad, 5 * NP);
        E.so.assign(cfg->nt, 0.f);
    }
    // model arrays start at zero, so the padding values of each element slot
    // (read only by the tail lanes, which write only padding) are zeros
    {
        float *model[6] = {c->lm, c->A, c->B, c->C, c->rhoinv, c->Q};
        const long long n[6] = {2 * NP, NP, NP, NP, NP, NP};
        for (int k = 0; k < 6 && !rc; ++k)
            if (model[k]) rc = zero(c, model[k], sizeof(float) * n[k]);
    }
    if (rc) {
        ses3d_hip_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

void ses3d_hip_destroy(ses3d_hip_context *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->release_runner();
    for (auto &a : c->allocs) (void)hipFree(a.first);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int ses3d_hip_set_model(ses3d_hip_context *c, const float *rhoinv, const float *mu, const float *lambda,
                        const float *A, const float *B, const float *C, const float *Q) {
    if (!c || !rhoinv || !mu || !lambda || !A || !B || !C) return fail(SES3D_HIP_EINVAL, "null model field");
    if (c->cfg.is_diss && !Q) return fail(SES3D_HIP_EINVAL, "is_diss=1 needs Q");
    const float *src[5] = {rhoinv, A, B, C, Q};
    float *dst[5] = {c->rhoinv, c->A, c->B, c->C, c->Q};
    for (int f = 0; f < 5; ++f)
        if (dst[f] && src[f]) TRY(upload_ref(c, dst[f], src[f], 1));
    TRY(upload_ref(c, c->lm, lambda, 2));       // (lambda, mu) pairs
    TRY(upload_ref(c, c->lm + 1, mu, 2));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->aniso = false;
    for (long long i = 0; i < c->NPTS && !c->aniso; ++i)
        if (A[i] != 0.0f || B[i] != 0.0f || C[i] != 0.0f) c->aniso = true;
    c->model_set = true;
    c->initialised = false;
    invalidate_checkpoints(c);
    return 0;
}

int ses3d_hip_set_source(ses3d_hip_context *c, int32_t ev, int32_t source_type, float xxs, float yys,
                         float zzs, const float *mom, const float *so) {
    TRY(check_ev(c, ev));
    if (!mom || !so) return fail(SES3D_HIP_EINVAL, "null source argument");
    if (!(source_type == 1 || source_type == 2 || source_type == 3 || source_type == 10))
        return fail(SES3D_HIP_EINVAL, "source_type must be 1, 2, 3 or 10");
    EventHost &E = c->evh[ev];
    E.source_type = source_type;
    E.xxs = xxs; E.yys = yys; E.zzs = zzs;
    for (int k = 0; k < 6; ++k) E.mom[k] = mom[k];
    E.so.assign(so, so + c->cfg.nt);
    E.have_source = true;
    c->initialised = false;
    return 0;
}

int ses3d_hip_set_receivers(ses3d_hip_context *c, int32_t ev, int32_t nr, const float *rec) {
    TRY(check_ev(c, ev));
    if (nr < 0 || (nr > 0 && !rec)) return fail(SES3D_HIP_EINVAL, "bad receiver list");
    c->evh[ev].rec.assign(rec, rec + 3 * (size_t)nr);
    c->initialised = false;
    return 0;
}

int ses3d_hip_set_adjoint_sources(ses3d_hip_context *c, int32_t ev, int32_t n, const float *loc, const float *stf) {
    TRY(check_ev(c, ev));
    if (n < 0 || (n > 0 && (!loc || !stf))) return fail(SES3D_HIP_EINVAL, "bad adjoint source list");
    c->evh[ev].ad_loc.assign(loc, loc + 3 * (size_t)n);
    c->evh[ev].ad_stf.assign(stf, stf + (size_t)n * c->cfg.nt * 3);
    c->initialised = false;
    return 0;
}

int ses3d_hip_init(ses3d_hip_context *c) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (!c->model_set) return fail(SES3D_HIP_ESTATE, "set_model must precede init");
    const auto &ms = c->ms;
    const int nt = c->cfg.nt;
    const int af = c->cfg.adjoint_flag;
    // small per-axis tables (allocated once)
    if (!c->Z) {
        TRY(dalloc(c, &c->Z, ms.z.size()));
        TRY(dalloc(c, &c->S, ms.sin_x.size()));
        TRY(dalloc(c, &c->COT, ms.cot_x.size()));
        TRY(dalloc(c, &c->IX, ms.IX.size()));
        TRY(dalloc(c, &c->IY, ms.IY.size()));
        TRY(dalloc(c, &c->IZ, ms.IZ.size()));
        TRY(dalloc(c, &c->LUT, ms.LUT.size()));
        TRY(dalloc(c, &c->w5, 5));
        TRY(dalloc(c, &c->PMLX, c->NX));
        TRY(dalloc(c, &c->PMLY, c->NY));
        TRY(dalloc(c, &c->PMLZ, c->NZ));
    }
    auto any_layer = [](const std::vector<uint16_t> &ids, int ne, std::vector<unsigned char> &out) {
        out.assign(ne, 0);
        for (int e = 0; e < ne; ++e)
            for (int q = 0; q < NN; ++q) out[e] |= ids[e * NN + q] != 0;
    };
    any_layer(ms.IX, c->NX, c->pmlx);
    any_layer(ms.IY, c->NY, c->pmly);
    any_layer(ms.IZ, c->NZ, c->pmlz);
    TRY(upload(c, c->PMLX, c->pmlx.data(), c->NX));
    TRY(upload(c, c->PMLY, c->pmly.data(), c->NY));
    TRY(upload(c, c->PMLZ, c->pmlz.data(), c->NZ));
    TRY(upload(c, c->Z, ms.z.data(), sizeof(float) * ms.z.size()));
    TRY(upload(c, c->S, ms.sin_x.data(), sizeof(float) * ms.sin_x.size()));
    TRY(upload(c, c->COT, ms.cot_x.data(), sizeof(float) * ms.cot_x.size()));
    TRY(upload(c, c->IX, ms.IX.data(), sizeof(uint16_t) * ms.IX.size()));
    TRY(upload(c, c->IY, ms.IY.data(), sizeof(uint16_t) * ms.IY.size()));
    TRY(upload(c, c->IZ, ms.IZ.data(), sizeof(uint16_t) * ms.IZ.size()));
    TRY(upload(c, c->LUT, ms.LUT.data(), sizeof(float) * ms.LUT.size()));
    TRY(upload(c, c->w5, ms.w, sizeof(float) * 5));
    MeshDev &M = c->mesh_h;
    M.NX = c->NX; M.NY = c->NY; M.NZ = c->NZ; M.NBX = c->NBX; M.NBY = c->NBY; M.NBZ = c->NBZ;
    M.NE = c->NE; M.NP = c->NP; M.NGB = c->NGB;
    M.KX = ms.KX; M.KY = ms.KY; M.KZ = ms.KZ;
    M.is_diss = c->cfg.is_diss;
    M.dt = c->cfg.dt; M.Jac = ms.Jac;
    M.sum_D_p = (c->cfg.D_p[0] + c->cfg.D_p[1]) + c->cfg.D_p[2];
    for (int k = 0; k < 3; ++k) {
        M.tau_p[k] = c->cfg.tau_p[k];
        M.Lk[k] = (c->cfg.dt * c->cfg.D_p[k]) / c->cfg.tau_p[k];   // L=dt*D_p(k)/tau_p(k), :338
    }
    std::memcpy(M.tab + 0, ms.D1, sizeof(ms.D1));
    std::memcpy(M.tab + 25, ms.D2, sizeof(ms.D2));
    std::memcpy(M.tab + 50, ms.D3, sizeof(ms.D3));
    std::memcpy(M.tab + 75, ms.WDL, sizeof(ms.WDL));
    std::memcpy(M.tab + 100, ms.DUM1, sizeof(ms.DUM1));
    std::memcpy(M.tab + 125, ms.DUM2, sizeof(ms.DUM2));
    std::memcpy(M.tab + 150, ms.DUM3, sizeof(ms.DUM3));
    ses3d::stored_degree_tables(c->hdeg, M.tab + 175, M.tab + 195);
    M.hdeg = c->hdeg;
    M.hnc = (c->hdeg + 1) * (c->hdeg + 1) * (c->hdeg + 1);
    M.hslot = hist_slot_floats(c);
    M.kf32 = c->kernel_bits == 32;
    M.nt = c->cfg.nt;
    M.lm = c->lm; M.A = c->A; M.B = c->B; M.C = c->C; M.Q = c->Q; M.rhoinv = c->rhoinv;
    M.MM = c->MM; M.Z = c->Z; M.S = c->S; M.COT = c->COT; M.IX = c->IX; M.IY = c->IY; M.IZ = c->IZ;
    M.LUT = c->LUT;
    M.PMLX = c->PMLX; M.PMLY = c->PMLY; M.PMLZ = c->PMLZ;
    {
        // reciprocal-based quotients: verified exhaustively on this device,
        // once per divisor in the process
        float bh[4] = {c->cfg.tau_p[0], c->cfg.tau_p[1], c->cfg.tau_p[2], 3.0f};
        float yh[4];
        for (int k = 0; k < 4; ++k) yh[k] = 1.0f / bh[k];
        unsigned long long bad[4];
        float bn[4], yn[4];
        int nn = 0, which[4];
        for (int k = 0; k < 4; ++k)
            if (!div_cached(c->device, bh[k], &bad[k])) {
                bn[nn] = bh[k];
                yn[nn] = yh[k];
                which[nn++] = k;
            }
        if (nn) {
            if (!c->div_buf) TRY(dalloc(c, &c->div_buf, 8));
            float *bd = (float *)c->div_buf, *yd = bd + 4;
            unsigned long long *badd = c->div_buf + 4;
            TRY(upload(c, bd, bn, sizeof(float) * nn));
            TRY(upload(c, yd, yn, sizeof(float) * nn));
            HIP_TRY(hipMemsetAsync(badd, 0, sizeof(unsigned long long) * 4, c->stream));
            hipLaunchKernelGGL(k_verify_div, dim3(8192), dim3(256), 0, c->stream, bd, yd, nn, badd);
            HIP_TRY(hipGetLastError());
            unsigned long long got[4];
            TRY(download(c, got, badd, sizeof(got)));
            for (int j = 0; j < nn; ++j) {
                bad[which[j]] = got[j];
                div_store(c->device, bn[j], got[j]);
            }
        }
        for (int k = 0; k < 3; ++k) M.rtau_p[k] = yh[k];
        M.r3 = yh[3];
        M.fast_tau = c->cfg.is_diss && bad[0] == 0 && bad[1] == 0 && bad[2] == 0;
        M.fast_3 = bad[3] == 0;
#ifdef SES3D_NO_RECIP_DIV
        M.fast_tau = M.fast_3 = 0;
#endif
        c->div_bad[0] = bad[0]; c->div_bad[1] = bad[1]; c->div_bad[2] = bad[2]; c->div_bad[3] = bad[3];
    }
    TRY(upload(c, c->mesh_d, &M, sizeof(MeshDev)));
    // mass matrix
    {
        const long long threads = node_blocks(c) * 64;
        hipLaunchKernelGGL(k_mass, dim3((unsigned)((threads + DEFAULT_NODE_THREADS - 1) / DEFAULT_NODE_THREADS)),
                           dim3(DEFAULT_NODE_THREADS), 0,
                           c->stream, c->mesh_d, c->MM, c->w5);
        HIP_TRY(hipGetLastError());
    }
    // events: zero fields, locate source / receivers / adjoint sources
    std::vector<RecDev> recs;
    std::vector<ForceDev> forces;                 // every event's sources, grouped by element
    std::vector<std::vector<int2>> force_elem(c->n_events);
    c->saving_vector.assign(nt + 1, 0);
    for (int ev = 0; ev < c->n_events; ++ev) {
        EventHost &E = c->evh[ev];
        TRY(zero(c, E.v, sizeof(float) * 3 * c->NGB));
        TRY(zero(c, E.st, sizeof(float) * state_floats(c)));
        if (E.grad) TRY(zero(c, E.grad, sizeof(double) * 5 * c->NP));
        TRY(upload(c, E.so_d, E.so.data(), sizeof(float) * nt));
        std::vector<float> md(9 * 125, 0.f);
        E.sloc = ses3d::SourceLoc();
        if ((af == 0 || af == 1) && E.have_source) {
            E.sloc = ses3d::locate_source(ms, E.xxs, E.yys, E.zzs);
            if (E.sloc.is_source && E.source_type == 10) ses3d::moment_weights(ms, E.sloc, E.mom, md.data());
        }
        TRY(upload(c, E.src_md, md.data(), sizeof(float) * md.size()));
        // receivers
        E.recs.clear();
        if (af == 0 || af == 1) E.recs = ses3d::locate_receivers(ms, E.rec.data(), (int)(E.rec.size() / 3));
        const int nr = (int)E.recs.size();
        if (nr > 0) {
            TRY(dfit(c, &E.seis, E.seis_cap, (size_t)nr * 3 * nt));
            TRY(dfit(c, &E.rec_w, E.rec_w_cap, (size_t)nr * 125));
            TRY(zero(c, E.seis, sizeof(float) * (size_t)nr * 3 * nt));
            std::vector<float> w((size_t)nr * 125);
            for (int r = 0; r < nr; ++r) {
                ses3d::point_weights(E.recs[r].xl, E.recs[r].yl, E.recs[r].zl, &w[(size_t)r * 125]);
                RecDev R;
                R.v = E.v; R.itp = c->it_d; R.ngb = c->NGB; R.nbx = c->NBX; R.nby = c->NBY;
                R.ex = E.recs[r].ex; R.ey = E.recs[r].ey; R.ez = E.recs[r].ez;
                R.nt = nt;
                R.w = E.rec_w + (size_t)r * 125;
                R.out = E.seis + (size_t)r * 3 * nt;
                recs.push_back(R);
            }
            TRY(upload(c, E.rec_w, w.data(), sizeof(float) * w.size()));
        }
        // point forces: forward single force or adjoint sources, grouped by element in order
        std::vector<std::pair<long long, std::pair<int, std::vector<float>>>> plist;   // elem, (mask, w)
        std::vector<std::vector<float>> pstf;
        if ((af == 0 || af == 1) && E.have_source && E.sloc.is_source && E.source_type < 10) {
            std::vector<float> w(125);
            ses3d::point_weights(E.sloc.xl, E.sloc.yl, E.sloc.zl, w.data());
            long long el = E.sloc.isx + (long long)c->NX * (E.sloc.isy + (long long)c->NY * E.sloc.isz);
            std::vector<float> stf(3 * (size_t)nt, 0.f);
            std::copy(E.so.begin(), E.so.end(), stf.begin() + (size_t)(E.source_type - 1) * nt);
            plist.push_back({el, {1 << (E.source_type - 1), w}});
            pstf.push_back(stf);
        }
        E.ads.clear();
        if (af == 2) {
            E.ads = ses3d::locate_adjoint_sources(ms, E.ad_loc.data(), (int)(E.ad_loc.size() / 3));
            for (auto &a : E.ads) {
                std::vector<float> w(125);
                ses3d::point_weights(a.xl, a.yl, a.zl, w.data());
                long long el = a.ex + (long long)c->NX * (a.ey + (long long)c->NY * a.ez);
                std::vector<float> stf(3 * (size_t)nt);
                for (int t = 0; t < nt; ++t)
                    for (int cc = 0; cc < 3; ++cc) stf[(size_t)cc * nt + t] = E.ad_stf[((size_t)a.index * nt + t) * 3 + cc];
                plist.push_back({el, {7, w}});
                pstf.push_back(stf);
            }
        }
        if (!plist.empty()) {
            const size_t np = plist.size();
            TRY(dfit(c, &E.force_w, E.force_w_cap, np * 125));
            TRY(dfit(c, &E.force_stf, E.force_stf_cap, np * 3 * (size_t)nt));
            std::vector<float> wall(np * 125), sall(np * 3 * (size_t)nt);
            for (size_t k = 0; k < np; ++k) {
                std::copy(plist[k].second.second.begin(), plist[k].second.second.end(), wall.begin() + k * 125);
                std::copy(pstf[k].begin(), pstf[k].end(), sall.begin() + k * 3 * (size_t)nt);
            }
            TRY(upload(c, E.force_w, wall.data(), sizeof(float) * wall.size()));
            TRY(upload(c, E.force_stf, sall.data(), sizeof(float) * sall.size()));
            // group by element keeping source order within each element
            std::map<long long, std::vector<size_t>> by_elem;
            for (size_t k = 0; k < np; ++k) by_elem[plist[k].first].push_back(k);
            force_elem[ev].assign((size_t)c->NE, int2{0, 0});
            for (auto &kv : by_elem) {
                force_elem[ev][(size_t)kv.first] = int2{(int)forces.size(), (int)kv.second.size()};
                for (size_t k : kv.second) {
                    ForceDev F;
                    F.comp_mask = plist[k].second.first;
                    F.w = E.force_w + k * 125;
                    F.stf = E.force_stf + k * 3 * (size_t)nt;
                    forces.push_back(F);
                }
            }
        }
        E.has_forces = !plist.empty();
        // sampled-step slots: adjoint_flag 1 stores, 2 reads the fwd step nt+1-it
        if (af == 1 || af == 2) {
            TRY(build_slots(c, E, false));
            E.hist = nullptr;
            TRY(upload_slots(c, std::vector<int>(nt + 1, -1)));
            for (int it = 1; it <= nt; ++it) c->saving_vector[it] = E.slot_h[it] >= 0 ? it : 0;
        }
    }
    // one element launch: the elements' workgroups of every event and one
    // workgroup per receiver, 128 to 512 lanes each, below 2^32 lanes
    if (((long long)c->NE * c->n_events + 4LL * ((long long)recs.size() + c->n_events)) * TPE >= (1LL << 32))
        return fail(SES3D_HIP_EINVAL, "events x elements plus receivers exceed one launch: split the events over contexts");
    c->nrec_total = (int)recs.size();
    c->recs_h = recs;
    if (!recs.empty()) {
        TRY(dfit(c, &c->recs_d, c->recs_cap, recs.size()));
        TRY(upload(c, c->recs_d, recs.data(), sizeof(RecDev) * recs.size()));
    }
    if (!forces.empty()) {
        TRY(dfit(c, &c->forces_d, c->forces_cap, forces.size()));
        TRY(upload(c, c->forces_d, forces.data(), sizeof(ForceDev) * forces.size()));
        for (int ev = 0; ev < c->n_events; ++ev) {
            EventHost &E = c->evh[ev];
            if (!E.has_forces) continue;
            TRY(dfit(c, &E.force_elem, E.force_elem_cap, (size_t)c->NE));
            TRY(upload(c, E.force_elem, force_elem[ev].data(), sizeof(int2) * (size_t)c->NE));
        }
    }
    TRY(refresh_events(c));
    c->drop_graphs();
    invalidate_checkpoints(c);
    c->it_host = -1;
    c->generation += 1;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->initialised = true;
    return 0;
}

static int run_checked(ses3d_hip_context *c, int32_t it_begin, int32_t n_steps, bool tune) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (!c->initialised) return fail(SES3D_HIP_ESTATE, "init must precede run");
    if (it_begin < 1 || n_steps < 0 || it_begin + n_steps - 1 > c->cfg.nt)
        return fail(SES3D_HIP_EINVAL, "step range outside 1..nt");
    if (tune && n_steps < N_TRIALS) return fail(SES3D_HIP_EINVAL, "autotune needs at least 6 steps");
    const int af = c->cfg.adjoint_flag;
    if (af == 1) invalidate_checkpoints(c);
    if (af == 1 || af == 2) TRY(ensure_full_history(c));
    return timed_run(c, it_begin, n_steps, af, af != 2, tune);
}

int ses3d_hip_run(ses3d_hip_context *c, int32_t it_begin, int32_t n_steps) {
    return run_checked(c, it_begin, n_steps, false);
}

int ses3d_hip_autotune(ses3d_hip_context *c, int32_t it_begin, int32_t n_steps) {
    return run_checked(c, it_begin, n_steps, true);
}

int ses3d_hip_set_schedule(ses3d_hip_context *c, const int32_t *sched) {
    if (!c || !sched) return fail(SES3D_HIP_EINVAL, "null argument");
    return set_schedule(c, sched);
}

int ses3d_hip_get_schedule(ses3d_hip_context *c, int32_t *sched, float *trial_ms) {
    if (!c || !sched) return fail(SES3D_HIP_EINVAL, "null argument");
    get_schedule(c, sched, trial_ms);
    return 0;
}

int ses3d_hip_set_streaming(ses3d_hip_context *c, int32_t mode) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    return set_streaming(c, mode);
}

int ses3d_hip_get_streaming(ses3d_hip_context *c, int32_t *mode, int32_t *in_use) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (mode) *mode = c->streaming;
    if (in_use) *in_use = use_streaming(c) ? 1 : 0;
    return 0;
}

int ses3d_hip_synchronize(ses3d_hip_context *c) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int ses3d_hip_get_nr(ses3d_hip_context *c, int32_t ev, int32_t *nr) {
    TRY(check_ev(c, ev));
    if (!nr) return fail(SES3D_HIP_EINVAL, "null nr");
    *nr = (int32_t)c->evh[ev].recs.size();
    return 0;
}

int ses3d_hip_get_receiver_index(ses3d_hip_context *c, int32_t ev, int32_t *idx) {
    TRY(check_ev(c, ev));
    if (!idx && !c->evh[ev].recs.empty()) return fail(SES3D_HIP_EINVAL, "null idx");
    for (size_t r = 0; r < c->evh[ev].recs.size(); ++r) idx[r] = c->evh[ev].recs[r].index;
    return 0;
}

int ses3d_hip_get_seismograms(ses3d_hip_context *c, int32_t ev, float *out) {
    TRY(check_ev(c, ev));
    EventHost &E = c->evh[ev];
    const int nr = (int)E.recs.size(), nt = c->cfg.nt;
    if (nr == 0) return 0;
    if (!out) return fail(SES3D_HIP_EINVAL, "null out");
    std::vector<float> tmp((size_t)nr * 3 * nt);
    TRY(download(c, tmp.data(), E.seis, sizeof(float) * tmp.size()));
    for (int r = 0; r < nr; ++r)
        for (int cc = 0; cc < 3; ++cc)
            std::memcpy(out + ((size_t)cc * nr + r) * nt, &tmp[((size_t)r * 3 + cc) * nt], sizeof(float) * nt);
    return 0;
}

int ses3d_hip_get_field(ses3d_hip_context *c, int32_t ev, const char *name, float *out) {
    TRY(check_ev(c, ev));
    if (!name || !out) return fail(SES3D_HIP_EINVAL, "null argument");
    EventHost &E = c->evh[ev];
    const std::string nm(name);
    const StateLayout SL(c->cfg.is_diss);
    static const char *dn[9] = {"dxux", "dyuy", "dzuz", "dxuy", "dyux", "dxuz", "dzux", "dyuz", "dzuy"};
    static const char *sn[9] = {"sxx", "syy", "szz", "sxy", "syx", "sxz", "szx", "syz", "szy"};
    static const char *mn[6] = {"Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"};
    // outside the absorbing layers the second member of each split-stress pair
    // is the first (only the first is stored there)
    static const int pair[9] = {EXX, EYY, EZZ, EXY, EXY, EXZ, EXZ, EYZ, EYZ};
    for (int k = 0; k < 9; ++k) {
        if (nm == dn[k]) {
            if (c->cfg.adjoint_flag == 0)
                return fail(SES3D_HIP_EINVAL, "with adjoint_flag=0 the displacement gradient is integrated only "
                                              "inside the absorbing layers (the only place it acts); use adjoint_flag 1 or 2");
            return download_ref(c, out, E.st + SL.d_off + c9_block_offset(k, 3 * StateLayout::PS), c9_stride(k),
                                SL.esb);
        }
        if (nm == sn[k]) {
            TRY(download_ref(c, out, E.st + c9_block_offset(k, SL.s_out), c9_stride(k), SL.esb));
            if (pair[k] != k) {
                std::vector<float> first(c->NPTS);
                TRY(download_ref(c, first.data(), E.st + c9_block_offset(pair[k], SL.s_out), c9_stride(pair[k]),
                                 SL.esb));
                std::vector<char> inner(c->NE);
                for (long long e = 0; e < c->NE; ++e) {
                    const int i = (int)(e % c->NX), j = (int)((e / c->NX) % c->NY), kk = (int)(e / ((long long)c->NX * c->NY));
                    inner[e] = !(c->pmlx[i] | c->pmly[j] | c->pmlz[kk]);
                }
                for (long long q = 0; q < c->NPTS; ++q)
                    if (inner[q % c->NE]) out[q] = first[q];
            }
            return 0;
        }
    }
    for (int k = 0; k < 6; ++k)
        if (nm == mn[k]) {
            if (!c->cfg.is_diss) return fail(SES3D_HIP_EINVAL, "memory variables exist only when is_diss=1");
            for (int mech = 0; mech < 3; ++mech) {
                const int i = k * 3 + mech;   // pair plane i / 2, member i % 2
                TRY(download_ref(c, out + mech * c->NPTS, E.st + SL.m_off + (i >> 1) * StateLayout::PS + (i & 1), 2,
                                 SL.esb));
            }
            return 0;
        }
    const int gx = 4 * c->NX + 1, gy = 4 * c->NY + 1, gz = 4 * c->NZ + 1;
    // global arrays: one plane (stride 1) or a member of the (vx, vy) pair plane
    auto global_out = [&](const float *dev, int stride) -> int {
        std::vector<float> g(c->NGB * stride);
        TRY(download(c, g.data(), dev, sizeof(float) * c->NGB * stride));
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x)
                    out[x + (long long)gx * (y + (long long)gy * z)] = g[host_gidx(c, x, y, z) * stride];
        return 0;
    };
    auto local_out = [&](const float *dev, int stride) -> int {
        std::vector<float> g(c->NGB * stride);
        TRY(download(c, g.data(), dev, sizeof(float) * c->NGB * stride));
        const long long NE = c->NE;
        for (int n = 0; n < 5; ++n)
            for (int m = 0; m < 5; ++m)
                for (int l = 0; l < 5; ++l)
                    for (int k = 0; k < c->NZ; ++k)
                        for (int j = 0; j < c->NY; ++j)
                            for (int i = 0; i < c->NX; ++i) {
                                long long e = i + (long long)c->NX * (j + (long long)c->NY * k);
                                out[e + NE * (l + 5 * (m + 5 * n))] = g[host_gidx(c, 4 * i + l, 4 * j + m, 4 * k + n) * stride];
                            }
        return 0;
    };
    if (nm == "vx_global") return global_out(E.v, 2);
    if (nm == "vy_global") return global_out(E.v + 1, 2);
    if (nm == "vz_global") return global_out(E.v + 2 * c->NGB, 1);
    if (nm == "MM_global") return global_out(c->MM, 1);
    // element-local velocity as the snapshots hold it: the local copy gathered at the
    // start of the last step, i.e. the velocity before that step's update.  Between
    // steps it equals a gather of the current global velocity of the previous step;
    // here it is the gather of the current global velocity (the state after the last
    // completed step), which is what the next step's gather will read.
    if (nm == "vx") return local_out(E.v, 2);
    if (nm == "vy") return local_out(E.v + 1, 2);
    if (nm == "vz") return local_out(E.v + 2 * c->NGB, 1);
    return fail(SES3D_HIP_EINVAL, "unknown field name " + nm);
}

int ses3d_hip_get_kernels(ses3d_hip_context *c, int32_t ev, double *g_rho, double *g_cp, double *g_csh,
                          double *g_csv, double *g_Q) {
    TRY(check_ev(c, ev));
    EventHost &E = c->evh[ev];
    if (!E.grad) return fail(SES3D_HIP_ESTATE, "kernels exist only for adjoint_flag=2 contexts");
    const long long NP = c->NP;
    const bool f32 = c->kernel_bits == 32;
    double *outs[5] = {g_rho, g_cp, g_csh, g_csv, g_Q};
    std::vector<float> tmp(f32 ? c->NPTS : 0);
    for (int k = 0; k < 5; ++k) {
        if (!outs[k]) continue;
        if (k == 4 && !c->cfg.is_diss) return fail(SES3D_HIP_EINVAL, "grad_Q_mu needs is_diss=1");
        // pair planes (rho, cp) (csh, csv), then Q_mu
        const long long off = k < 4 ? (k / 2) * 2 * NP + (k % 2) : 4 * NP;
        const int stride = k < 4 ? 2 : 1;
        if (f32) {
            TRY(download_ref(c, tmp.data(), (const float *)E.grad + off, stride));
            for (long long i = 0; i < c->NPTS; ++i) outs[k][i] = (double)tmp[i];
        } else {
            TRY(download_ref(c, outs[k], E.grad + off, stride));
        }
    }
    if (g_Q && c->cfg.is_diss) {
        // normalisation at it==nt (SOURCE/ses3d_grad_Q.f90:151-153): 2 mu K / Q,
        // in the accumulators' type
        std::vector<float> mur(c->NPTS), qr(c->NPTS);
        TRY(download_ref(c, mur.data(), (const float *)c->lm + 1, 2));
        TRY(download_ref(c, qr.data(), (const float *)c->Q, 1));
        if (f32)
            for (long long i = 0; i < c->NPTS; ++i)
                g_Q[i] = (double)(((2.0f * mur[i]) * (float)g_Q[i]) / qr[i]);
        else
            for (long long i = 0; i < c->NPTS; ++i) g_Q[i] = ((2.0 * (double)mur[i]) * g_Q[i]) / (double)qr[i];
    }
    return 0;
}

int ses3d_hip_set_model_device(ses3d_hip_context *c, const float *rhoinv, const float *mu, const float *lambda,
                               const float *A, const float *B, const float *C, const float *Q) {
    if (!c || !rhoinv || !mu || !lambda || !A || !B || !C) return fail(SES3D_HIP_EINVAL, "null model field");
    if (c->cfg.is_diss && !Q) return fail(SES3D_HIP_EINVAL, "is_diss=1 needs Q");
    const long long n = c->NPTS;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    const float *src[5] = {rhoinv, A, B, C, Q};
    float *dst[5] = {c->rhoinv, c->A, c->B, c->C, c->Q};
    for (int f = 0; f < 5; ++f)
        if (dst[f] && src[f])
            hipLaunchKernelGGL(k_from_ref<float>, grid, block, 0, c->stream, dst[f], src[f], c->NE, 1, (long long)EST);
    hipLaunchKernelGGL(k_from_ref<float>, grid, block, 0, c->stream, c->lm, lambda, c->NE, 2, 2LL * EST);   // (lambda, mu) pairs
    hipLaunchKernelGGL(k_from_ref<float>, grid, block, 0, c->stream, c->lm + 1, mu, c->NE, 2, 2LL * EST);
    HIP_TRY(hipGetLastError());
    // anisotropy test on the device (A, B, C all zero: isotropic kernels)
    if (!c->flag_d) TRY(dalloc(c, &c->flag_d, 1));
    HIP_TRY(hipMemsetAsync(c->flag_d, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_any_nonzero, grid, block, 0, c->stream, A, B, C, n, c->flag_d);
    HIP_TRY(hipGetLastError());
    int any = 0;
    TRY(download(c, &any, c->flag_d, sizeof(int)));   // also: the caller's buffers are no longer read
    c->aniso = any != 0;
    c->model_set = true;
    c->initialised = false;
    invalidate_checkpoints(c);
    return 0;
}

int ses3d_hip_get_kernels_device(ses3d_hip_context *c, int32_t ev, double *g_rho, double *g_cp, double *g_csh,
                                 double *g_csv, double *g_Q) {
    TRY(check_ev(c, ev));
    EventHost &E = c->evh[ev];
    if (!E.grad) return fail(SES3D_HIP_ESTATE, "kernels exist only for adjoint_flag=2 contexts");
    if (g_Q && !c->cfg.is_diss) return fail(SES3D_HIP_EINVAL, "grad_Q_mu needs is_diss=1");
    const long long NP = c->NP, n = c->NPTS;
    const bool f32 = c->kernel_bits == 32;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    double *outs[5] = {g_rho, g_cp, g_csh, g_csv, g_Q};
    for (int k = 0; k < 5; ++k) {
        if (!outs[k]) continue;
        // pair planes (rho, cp) (csh, csv), then Q_mu
        const long long off = k < 4 ? (k / 2) * 2 * NP + (k % 2) : 4 * NP;
        const int stride = k < 4 ? 2 : 1;
        if (f32)
            hipLaunchKernelGGL(k_to_ref_widen, grid, block, 0, c->stream, outs[k], (const float *)E.grad + off, c->NE,
                               stride);
        else
            hipLaunchKernelGGL(k_to_ref<double>, grid, block, 0, c->stream, outs[k], (const double *)E.grad + off,
                               c->NE, stride, (long long)EST * stride);
    }
    if (g_Q) hipLaunchKernelGGL(k_q_normalise, grid, block, 0, c->stream, g_Q, c->lm, c->Q, c->NE, f32 ? 1 : 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int ses3d_hip_get_history(ses3d_hip_context *c, int32_t ev, int32_t step_, float *out9) {
    TRY(check_ev(c, ev));
    if (!out9) return fail(SES3D_HIP_EINVAL, "null out9");
    EventHost &E = c->evh[ev];
    if (!E.hist_own) return fail(SES3D_HIP_ESTATE, "no stored forward history (run an adjoint_flag=1 context first)");
    if (step_ < 1 || step_ > c->cfg.nt || E.slot_h.empty() || E.slot_h[step_] < 0)
        return fail(SES3D_HIP_EINVAL, "step is not a stored forward step");
    const long long NP = c->NP;
    const float *src = E.hist_own + (long long)E.slot_h[step_] * hist_slot_floats(c);
    if (c->hdeg < 4) {
        // device [field][element][node] -> reference (0:nx,0:ny,0:nz,0:deg,0:deg,0:deg) per field
        const long long NE = c->NE, NC = hist_slot_floats(c) / (9 * NE);
        std::vector<float> tmp(9 * NE * NC);
        TRY(download(c, tmp.data(), src, sizeof(float) * tmp.size()));
        for (long long k = 0; k < 9; ++k)
            for (long long e = 0; e < NE; ++e)
                for (long long i = 0; i < NC; ++i) out9[(k * NC + i) * NE + e] = tmp[(k * NE + e) * NC + i];
        return 0;
    }
    // pair planes (vx,vy) (vz,exx) (eyy,ezz) (exy,exz), then eyz
    for (int k = 0; k < 9; ++k)
        TRY(download_ref(c, out9 + k * c->NPTS, k < 8 ? src + (k / 2) * 2 * NP + (k % 2) : src + 8 * NP, k < 8 ? 2 : 1));
    return 0;
}

int ses3d_hip_put_history(ses3d_hip_context *c, int32_t ev, int32_t step_, const float *in9) {
    TRY(check_ev(c, ev));
    if (!in9) return fail(SES3D_HIP_EINVAL, "null in9");
    if (!c->initialised) return fail(SES3D_HIP_ESTATE, "init must precede put_history");
    if (c->cfg.adjoint_flag != 2) return fail(SES3D_HIP_ESTATE, "put_history needs an adjoint_flag=2 context");
    TRY(ensure_full_history(c));
    EventHost &E = c->evh[ev];
    if (step_ < 1 || step_ > c->cfg.nt || E.slot_h.empty() || E.slot_h[step_] < 0)
        return fail(SES3D_HIP_EINVAL, "step is not a sampled forward step");
    const long long NP = c->NP;
    float *dst = E.hist_own + (long long)E.slot_h[step_] * hist_slot_floats(c);
    if (c->hdeg < 4) {
        const long long NE = c->NE, NC = hist_slot_floats(c) / (9 * NE);
        std::vector<float> tmp(9 * NE * NC);
        for (long long k = 0; k < 9; ++k)
            for (long long e = 0; e < NE; ++e)
                for (long long i = 0; i < NC; ++i) tmp[(k * NE + e) * NC + i] = in9[(k * NC + i) * NE + e];
        TRY(upload(c, dst, tmp.data(), sizeof(float) * tmp.size()));
        return 0;
    }
    // pair planes (vx,vy) (vz,exx) (eyy,ezz) (exy,exz), then eyz
    for (int k = 0; k < 9; ++k)
        TRY(upload_ref(c, k < 8 ? dst + (k / 2) * 2 * NP + (k % 2) : dst + 8 * NP, in9 + k * c->NPTS, k < 8 ? 2 : 1));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int ses3d_hip_set_history_degree(ses3d_hip_context *c, int32_t deg) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (deg < 0 || deg > 4) return fail(SES3D_HIP_EINVAL, "stored-field degree must be 0..4");
    if (deg == c->hdeg) return 0;
    // buffers of the other layout are released; init rebuilds the tables
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &E : c->evh) {
        if (E.hist_own) dfree(c, E.hist_own);
        if (E.seg) dfree(c, E.seg);
        E.hist_own = E.seg = E.hist = nullptr;
        E.seg_slots = 0;
    }
    c->hdeg = deg;
    c->initialised = false;
    invalidate_checkpoints(c);
    return 0;
}

int ses3d_hip_set_kernel_precision(ses3d_hip_context *c, int32_t bits) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (bits != 32 && bits != 64) return fail(SES3D_HIP_EINVAL, "kernel precision must be 32 or 64 bits");
    if (bits == c->kernel_bits) return 0;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->kernel_bits = bits;
    c->initialised = false;   // init zeroes the accumulators and passes the type to the kernels
    return 0;
}

int ses3d_hip_history_size(ses3d_hip_context *c, int64_t *n) {
    if (!c || !n) return fail(SES3D_HIP_EINVAL, "null argument");
    *n = c->hdeg < 4 ? hist_slot_floats(c) : 9 * c->NPTS;   // host (reference-layout) floats
    return 0;
}

int ses3d_hip_get_saving_vector(ses3d_hip_context *c, int32_t *sv) {
    if (!c || !sv) return fail(SES3D_HIP_EINVAL, "null argument");
    for (int it = 1; it <= c->cfg.nt; ++it) sv[it - 1] = c->saving_vector.empty() ? 0 : c->saving_vector[it];
    return 0;
}

int ses3d_hip_forward_checkpointed(ses3d_hip_context *c, int32_t k) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (!c->initialised) return fail(SES3D_HIP_ESTATE, "init must precede forward_checkpointed");
    if (c->cfg.adjoint_flag != 1) return fail(SES3D_HIP_ESTATE, "forward_checkpointed needs an adjoint_flag=1 context");
    if (k < 1) return fail(SES3D_HIP_EINVAL, "checkpoint interval must be >= 1");
    const int nt = c->cfg.nt;
    const int nseg = (nt + k - 1) / k;
    const int keep = std::min(c->ckpt_keep, nseg);
    const int first_kept = nseg - keep;   // the first segment whose sampled steps are kept
    const size_t NP = (size_t)c->NP, NGB = (size_t)c->NGB;
    const size_t nst = state_floats(c);
    TRY(ensure_checkpoints(c, k, keep));
    // the forward pass keeps the displacement gradient (mode 1) but stores no sampled step
    for (auto &E : c->evh) E.hist = nullptr;
    TRY(upload_slots(c, std::vector<int>(nt + 1, -1)));
    TRY(refresh_events(c));
    c->ckpt_k = k;
    c->ckpt_keep_used = keep;
    c->last_seg_ready = false;
    for (int sgi = 0; sgi < nseg; ++sgi) {
        if (sgi == first_kept) {
            // the trailing segments store their sampled steps right away: the
            // adjoint starts with them and need not replay them
            for (auto &E : c->evh) E.hist = E.seg;
            TRY(upload_slots(c, segment_table(c->evh[0], nt, sgi * k + 1, nt)));
            TRY(refresh_events(c));
        }
        // restart state of every segment (a kept one is replayed only by a
        // second adjoint, after the first overwrote the kept steps)
        for (auto &E : c->evh) {
            float *dst = E.ckpt[sgi];
            HIP_TRY(hipMemcpyAsync(dst, E.v, sizeof(float) * 3 * NGB, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(dst + 3 * NGB, E.st, sizeof(float) * nst, hipMemcpyDeviceToDevice, c->stream));
        }
        const int s0 = sgi * k + 1, s1 = std::min((sgi + 1) * k, nt);
        TRY(run_steps(c, s0, s1 - s0 + 1, 1, true));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ckpt_ready = true;
    c->last_seg_ready = true;
    return 0;
}

int ses3d_hip_set_checkpoint_keep(ses3d_hip_context *c, int32_t segments) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (segments < 1) return fail(SES3D_HIP_EINVAL, "kept segments must be >= 1");
    c->ckpt_keep = segments;
    return 0;
}

int ses3d_hip_prepare_checkpoints(ses3d_hip_context *c, int32_t k) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (!c->initialised) return fail(SES3D_HIP_ESTATE, "init must precede prepare_checkpoints");
    if (c->cfg.adjoint_flag != 1) return fail(SES3D_HIP_ESTATE, "prepare_checkpoints needs an adjoint_flag=1 context");
    if (k < 1) return fail(SES3D_HIP_EINVAL, "checkpoint interval must be >= 1");
    const int nseg = (c->cfg.nt + k - 1) / k;
    TRY(ensure_checkpoints(c, k, std::min(c->ckpt_keep, nseg)));
    // the buffers are cleared below: a previous checkpointed forward is gone
    c->ckpt_ready = false;
    c->last_seg_ready = false;
    // first touch of every page here rather than in the forward pass
    const size_t NP = (size_t)c->NP, NGB = (size_t)c->NGB;
    const size_t ck = 3 * NGB + state_floats(c);
    for (auto &E : c->evh) {
        for (float *q : E.ckpt) HIP_TRY(hipMemsetAsync(q, 0, sizeof(float) * ck, c->stream));
        if (E.seg && E.seg_slots > 0)
            HIP_TRY(hipMemsetAsync(E.seg, 0, sizeof(float) * (size_t)E.seg_slots * hist_slot_floats(c), c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int ses3d_hip_adjoint_checkpointed(ses3d_hip_context *a, ses3d_hip_context *f) {
    if (!a || !f) return fail(SES3D_HIP_EINVAL, "null context");
    if (!a->initialised) return fail(SES3D_HIP_ESTATE, "init the adjoint context first");
    if (a->cfg.adjoint_flag != 2) return fail(SES3D_HIP_ESTATE, "adjoint context needs adjoint_flag=2");
    if (!f->initialised || f->cfg.adjoint_flag != 1)
        return fail(SES3D_HIP_ESTATE, "the forward context must be an initialised adjoint_flag=1 context");
    if (!f->ckpt_ready)
        return fail(SES3D_HIP_ESTATE, "no current checkpointed forward: run forward_checkpointed on the forward "
                                      "context after its last set_model, init or run");
    if (a->cfg.nt != f->cfg.nt || a->NE != f->NE || a->n_events != f->n_events || a->cfg.samp_ad != f->cfg.samp_ad)
        return fail(SES3D_HIP_EINVAL, "forward and adjoint contexts differ in mesh, nt, samp_ad or event count");
    if (a->hdeg != f->hdeg) return fail(SES3D_HIP_EINVAL, "forward and adjoint contexts store fields of different degree");
    const int nt = f->cfg.nt, k = f->ckpt_k;
    const int nseg = (nt + k - 1) / k;
    const int first_kept = nseg - f->ckpt_keep_used;
    const size_t NP = (size_t)f->NP, NGB = (size_t)f->NGB;
    const size_t nst = state_floats(f);
    TRY(ensure_segment_buffer(f, k, f->ckpt_keep_used));
    for (int sgi = nseg - 1; sgi >= 0; --sgi) {
        const int s0 = sgi * k + 1, s1 = std::min((sgi + 1) * k, nt);
        // a kept trailing segment's steps were stored by forward_checkpointed
        // (first_kept*k+1..nt in the segment buffer); the first replay below
        // overwrites them
        const bool kept = sgi >= first_kept && f->last_seg_ready;
        if (!kept) {
            // restore the forward state at the start of the segment, replay it into the buffer
            f->last_seg_ready = false;
            for (int ev = 0; ev < f->n_events; ++ev) {
                EventHost &E = f->evh[ev];
                const float *src = E.ckpt[sgi];
                HIP_TRY(hipMemcpyAsync(E.v, src, sizeof(float) * 3 * NGB, hipMemcpyDeviceToDevice, f->stream));
                HIP_TRY(hipMemcpyAsync(E.st, src + 3 * NGB, sizeof(float) * nst, hipMemcpyDeviceToDevice, f->stream));
                E.hist = E.seg;
            }
            TRY(upload_slots(f, segment_table(f->evh[0], nt, s0, s1)));
            TRY(refresh_events(f));
            TRY(run_steps(f, s0, s1 - s0 + 1, 1, false));
            HIP_TRY(hipStreamSynchronize(f->stream));
        }
        // adjoint steps paired with this segment's forward steps, through the
        // slots of those steps in the segment buffer
        const int a0 = nt + 1 - s1, a1 = nt + 1 - s0;
        {
            const std::vector<int> seg_tab =
                segment_table(f->evh[0], nt, kept ? first_kept * k + 1 : s0, kept ? nt : s1);
            std::vector<int> tab(nt + 1, -1);
            for (int it = a0; it <= a1; ++it) tab[it] = seg_tab[nt + 1 - it];
            for (int ev = 0; ev < a->n_events; ++ev) a->evh[ev].hist = f->evh[ev].seg;
            TRY(upload_slots(a, tab));
        }
        TRY(refresh_events(a));
        TRY(run_steps(a, a0, a1 - a0 + 1, 2, false));
        HIP_TRY(hipStreamSynchronize(a->stream));
    }
    for (auto &A : a->evh) A.hist = A.hist_own;
    TRY(refresh_events(a));
    return 0;
}

int ses3d_hip_memory_bytes(ses3d_hip_context *c, int64_t *bytes) {
    if (!c || !bytes) return fail(SES3D_HIP_EINVAL, "null argument");
    *bytes = c->bytes;
    return 0;
}

int ses3d_hip_get_info(ses3d_hip_context *c, int64_t *out, int32_t n) {
    if (!c || !out) return fail(SES3D_HIP_EINVAL, "null argument");
    const int64_t v[8] = {c->mesh_h.fast_tau, c->mesh_h.fast_3, (int64_t)c->div_bad[0], (int64_t)c->div_bad[1],
                          (int64_t)c->div_bad[2], (int64_t)c->div_bad[3], c->aniso ? 1 : 0, c->graph_steps};
    for (int k = 0; k < n && k < 8; ++k) out[k] = v[k];
    return 0;
}

int ses3d_hip_set_graph_steps(ses3d_hip_context *c, int32_t steps) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (steps < 0) return fail(SES3D_HIP_EINVAL, "graph step count must be >= 0");
    c->drop_graphs();
    c->graph_steps = steps;
    return 0;
}

int ses3d_hip_set_profiling(ses3d_hip_context *c, int32_t on) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    c->profiling = on != 0;
    return 0;
}

int ses3d_hip_get_timing(ses3d_hip_context *c, double *total_ms, double *element_ms, double *node_ms,
                         int32_t *n_element_launches) {
    if (!c) return fail(SES3D_HIP_EINVAL, "null context");
    if (total_ms) *total_ms = c->t_total;
    if (element_ms) *element_ms = c->t_elem;
    if (node_ms) *node_ms = c->t_node;
    if (n_element_launches) *n_element_launches = c->n_elem;
    return 0;
}

// ---------------------------------------------------------------------------
// groups of sections
// ---------------------------------------------------------------------------
int ses3d_hip_group_create(ses3d_hip_context *const *ctxs, int32_t n, ses3d_hip_group **out) {
    if (!ctxs || !out) return fail(SES3D_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (n < 1) return fail(SES3D_HIP_EINVAL, "a group needs at least one context");
    for (int k = 0; k < n; ++k) {
        const ses3d_hip_context *c = ctxs[k];
        if (!c) return fail(SES3D_HIP_EINVAL, "null context in group");
        if (!c->initialised) return fail(SES3D_HIP_ESTATE, "every member must be initialised before grouping");
        for (int j = 0; j < k; ++j)
            if (ctxs[j] == c) return fail(SES3D_HIP_EINVAL, "a context appears twice in the group");
        if (c->device != ctxs[0]->device) return fail(SES3D_HIP_EINVAL, "members on different devices");
        if ((c->cfg.is_diss != 0) != (ctxs[0]->cfg.is_diss != 0))
            return fail(SES3D_HIP_EINVAL, "members must all be elastic or all anelastic");
        if (c->cfg.adjoint_flag != ctxs[0]->cfg.adjoint_flag)
            return fail(SES3D_HIP_EINVAL, "members must share adjoint_flag");
    }
    auto *g = new ses3d_hip_group();
    g->ctx.assign(ctxs, ctxs + n);
    g->diss = ctxs[0]->cfg.is_diss ? 1 : 0;
    g->af = ctxs[0]->cfg.adjoint_flag;
    for (auto *c : g->ctx) g->aniso |= c->aniso ? 1 : 0;
    int rc = 0;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(SES3D_HIP_EDEVICE, "hipStreamCreate failed");
    if (!rc && hipMalloc((void **)&g->secs_d, sizeof(SectionDev) * n) != hipSuccess)
        rc = fail(SES3D_HIP_ENOMEM, "hipMalloc of the group launch table failed");
    if (!rc) rc = g->build_table();
    if (rc) {
        ses3d_hip_group_destroy(g);
        return rc;
    }
    *out = g;
    return 0;
}

void ses3d_hip_group_destroy(ses3d_hip_group *g) {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    g->release_runner();
    if (g->secs_d) (void)hipFree(g->secs_d);
    if (g->recs_d) (void)hipFree(g->recs_d);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

static int group_run_checked(ses3d_hip_group *g, int32_t it_begin, int32_t n_steps, bool tune) {
    if (!g) return fail(SES3D_HIP_EINVAL, "null group");
    if (it_begin < 1 || n_steps < 0) return fail(SES3D_HIP_EINVAL, "step range outside 1..nt");
    if (tune && n_steps < N_TRIALS) return fail(SES3D_HIP_EINVAL, "autotune needs at least 6 steps");
    int aniso = 0;
    for (auto *c : g->ctx) {
        if (!c->initialised) return fail(SES3D_HIP_ESTATE, "every member must be initialised");
        if (it_begin + n_steps - 1 > c->cfg.nt) return fail(SES3D_HIP_EINVAL, "step range outside 1..nt of a member");
        aniso |= c->aniso ? 1 : 0;
    }
    for (auto *c : g->ctx) {
        if (g->af == 1) invalidate_checkpoints(c);
        if (g->af == 1 || g->af == 2) TRY(ensure_full_history(c));
        HIP_TRY(hipStreamSynchronize(c->stream));   // member work issued before this run
    }
    bool stale = aniso != g->aniso;
    for (size_t k = 0; k < g->ctx.size(); ++k) stale = stale || g->gens[k] != g->ctx[k]->generation;
    g->aniso = aniso;
    if (stale) TRY(g->schedule_changed());
    g->it_host = -1;   // members may have stepped on their own since the last group run
    const int rc = timed_run(g, it_begin, n_steps, g->af, g->af != 2, tune);
    // later work on a member's own stream starts after this run
    hipEvent_t done = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    hipError_t e = hipEventRecord(done, g->stream);
    for (auto *c : g->ctx) {
        if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, done, 0);
        c->it_host = rc ? -1 : it_begin + n_steps;
    }
    (void)hipEventDestroy(done);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SES3D_HIP_EDEVICE, std::string("group stream ordering: ") + hipGetErrorString(e));
    return 0;
}

int ses3d_hip_group_run(ses3d_hip_group *g, int32_t it_begin, int32_t n_steps) {
    return group_run_checked(g, it_begin, n_steps, false);
}

int ses3d_hip_group_autotune(ses3d_hip_group *g, int32_t it_begin, int32_t n_steps) {
    return group_run_checked(g, it_begin, n_steps, true);
}

int ses3d_hip_group_synchronize(ses3d_hip_group *g) {
    if (!g) return fail(SES3D_HIP_EINVAL, "null group");
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int ses3d_hip_group_set_schedule(ses3d_hip_group *g, const int32_t *sched) {
    if (!g || !sched) return fail(SES3D_HIP_EINVAL, "null argument");
    return set_schedule(g, sched);
}

int ses3d_hip_group_get_schedule(ses3d_hip_group *g, int32_t *sched, float *trial_ms) {
    if (!g || !sched) return fail(SES3D_HIP_EINVAL, "null argument");
    get_schedule(g, sched, trial_ms);
    return 0;
}

int ses3d_hip_group_set_streaming(ses3d_hip_group *g, int32_t mode) {
    if (!g) return fail(SES3D_HIP_EINVAL, "null group");
    return set_streaming(g, mode);
}

int ses3d_hip_group_get_streaming(ses3d_hip_group *g, int32_t *mode, int32_t *in_use) {
    if (!g) return fail(SES3D_HIP_EINVAL, "null group");
    if (mode) *mode = g->streaming;
    if (in_use) *in_use = use_streaming(g) ? 1 : 0;
    return 0;
}

int ses3d_hip_group_set_graph_steps(ses3d_hip_group *g, int32_t steps) {
    if (!g) return fail(SES3D_HIP_EINVAL, "null group");
    if (steps < 0) return fail(SES3D_HIP_EINVAL, "graph step count must be >= 0");
    g->drop_graphs();
    g->graph_steps = steps;
    return 0;
}

int ses3d_hip_group_set_profiling(ses3d_hip_group *g, int32_t on) {
    if (!g) return fail(SES3D_HIP_EINVAL, "null group");
    g->profiling = on != 0;
    return 0;
}

int ses3d_hip_group_get_timing(ses3d_hip_group *g, double *total_ms, double *element_ms, double *node_ms,
                               int32_t *n_element_launches) {
    if (!g) return fail(SES3D_HIP_EINVAL, "null group");
    if (total_ms) *total_ms = g->t_total;
    if (element_ms) *element_ms = g->t_elem;
    if (node_ms) *node_ms = g->t_node;
    if (n_element_launches) *n_element_launches = g->n_elem;
    return 0;
}

int ses3d_hip_project(ses3d_hip_context *c, const float *field, int32_t nbx, const float *bx_deg, int32_t nby,
                      const float *by_deg, int32_t nbz, const float *bz_km, float *out) {
    if (!c || !field || !bx_deg || !by_deg || !bz_km || !out) return fail(SES3D_HIP_EINVAL, "null argument");
    ProjPlan P;
    TRY(plan_projection(c, nbx, bx_deg, nby, by_deg, nbz, bz_km, P));
    // one device buffer for the field, the tables and the output (freed before return)
    char *buf = nullptr;
    HIP_TRY(hipMalloc((void **)&buf, P.bytes));
    const int rc = run_projection(c, P, field, nbx, nby, nbz, out, buf);
    (void)hipFree(buf);
    return rc;
}

int ses3d_hip_project_noalloc(ses3d_hip_context *c, const float *field, int32_t nbx, const float *bx_deg,
                              int32_t nby, const float *by_deg, int32_t nbz, const float *bz_km, float *out) {
    if (!c || !field || !bx_deg || !by_deg || !bz_km || !out) return fail(SES3D_HIP_EINVAL, "null argument");
    ProjPlan P;
    TRY(plan_projection(c, nbx, bx_deg, nby, by_deg, nbz, bz_km, P));
    if (P.bytes > (size_t)SES3D_NOALLOC_SCRATCH_BYTES)
        return fail(SES3D_HIP_EINVAL, "projection needs " + std::to_string(P.bytes) +
                                          " bytes of scratch (field + block tables + output); the no-allocation "
                                          "entry point holds at most " +
                                          std::to_string((long long)SES3D_NOALLOC_SCRATCH_BYTES) + " bytes");
    char *buf = (char *)scratch_address(c->device);
    if (!buf) return fail(SES3D_HIP_ESTATE, "no-allocation scratch was not set up on this context's device");
    return run_projection(c, P, field, nbx, nby, nbz, out, buf);
}

}  // extern "C"
