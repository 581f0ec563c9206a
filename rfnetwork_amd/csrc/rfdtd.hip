// rfdtd: FDTD time stepping for rfnetwork's field solver on MI355X (gfx950), behind the C ABI of
// include/rfdtd.h.
//
// Replaces the reference's GPU entry point (rfnetwork/core/src/cuda_binding.cpp:20-68 ->
// rfnetwork/core/src/solver.cu:546-1143) and follows the update semantics of the reference's CPU
// solver (rfnetwork/core/src/solver.cpp:848-1163):
//   * E half step: six split updates s = Ca*s + Cb*(h+ - h) with the upper-edge guards of the
//     GPU packing, then totals (solver.cu:152-190, solver.cpp:851-960);
//   * E soft sources / probes, in list order per cell (solver.cpp:963-976);
//   * H half step: s = Da*s + (Db2*e - Db1*e-) -- the CPU association -- then totals
//     (solver.cpp:992-1101);
//   * H probes (solver.cpp:1104-1116);
//   * monitors sampled after both half steps (solver.cpp:1130-1161): plane copies, or running DFT
//     values[f] += field * phase[m, f] in complex arithmetic (solver.cpp:1194-1225).
// All device code is compiled with -ffp-contract=off, so each product and sum is rounded on its
// own, exactly as in the reference's x86 build: for the same inputs the float results equal the CPU
// solver's bit for bit.
//
// Design for gfx950 (not a translation of the CUDA kernels):
//   * a session holds one or more grids ("members"), each advanced for one or more excitations
//     ("instances" sharing the member's coefficients); one launch per phase covers every instance
//     of every member, so a batch of uneven grids costs the launches of a single grid;
//   * separate E and H kernels: one thread owns a (y,z) column of a block of x planes and marches
//     up in x; the x+1 (E update) and x-1 (H update) neighbours are carried in registers, the y/z
//     neighbours come from the lanes of the same or the next wavefront (L1/L2 hits), so each array
//     is read from HBM once; where Nz allows, a thread owns 2 or 4 consecutive z cells and moves
//     every array as one 8- or 16-byte vector per lane;
//   * the two-step pass (k_step2) advances a tile through steps n and n + 1 in one launch: a
//     workgroup computes its tile plus a two-cell halo, exchanges y/z neighbours through LDS and
//     marches x with the second step one plane behind the first, so a cell's fields cross HBM once
//     per two steps (the halo recomputed, not re-read from another workgroup's results);
//   * lanes run along z, the contiguous axis: every load and store is a coalesced 256-byte row;
//   * workgroups are renumbered so that consecutive tiles land on the same XCD and share its L2:
//     neighbouring (y,z) tiles share y-neighbour rows, and the instances of one member -- ordered
//     innermost -- share the coefficient reads of the same cells.  The tiled passes always do so;
//     the separate kernels can instead take workgroups in launch order, which streams faster on
//     large grids (the autotuner chooses);
//   * the 30 coefficient arrays are deduplicated bit-exactly, on the device (hash, radix sort, runs
//     of equal tuples; a host pass on request), into per-cell 32-bit entry indices plus tables of
//     distinct E (12 values) and H (18 values) tuples: a step moves 8 B of coefficient index per
//     cell instead of 120 B wherever materials repeat;
//   * a chunk of steps is captured once into a hipGraph and replayed; the kernels read the step
//     number from a device word set before each replay.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../include/rfdtd.h"

namespace {

thread_local std::string g_last_error;

struct RfdtdError {
    int code;
    std::string msg;
};

#define RF_THROW(code, msg) throw RfdtdError{code, (msg)}
#define HIP_OK(expr)                                                                                         \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            throw RfdtdError{RFDTD_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)};              \
    } while (0)

enum { EX = 0, EY = 1, EZ = 2, HX = 3, HY = 4, HZ = 5 };

// Global address space qualifier for device code (the host pass parses device functions too).
#if defined(__HIP_DEVICE_COMPILE__)
#define GLOBAL_AS __attribute__((address_space(1)))
#define CONST_AS __attribute__((address_space(4)))  // read-only for the whole kernel: scalar loads
#else
#define GLOBAL_AS
#define CONST_AS
#endif

// ------------------------------------------------------------------------------------------------
// device side
// ------------------------------------------------------------------------------------------------

template <typename T>
struct Fields {
    T *ex_y, *ex_z, *ey_z, *ey_x, *ez_x, *ez_y;
    T *hx_y, *hx_z, *hy_z, *hy_x, *hz_x, *hz_y;
    T *ex, *ey, *ez, *hx, *hy, *hz;
};

template <typename T>
struct alignas(4 * sizeof(T)) Vec4 {
    T v[4];
};

// Load / store at a 32-bit byte offset from a wave-uniform base: lowers to the scalar-base +
// 32-bit-vector-offset form of global_load / global_store (no 64-bit address per lane).
template <class T>
__device__ __forceinline__ T ldg(const GLOBAL_AS T* base, uint32_t boff) {
    return *(const GLOBAL_AS T*)((const GLOBAL_AS char*)base + boff);
}
template <class T>
__device__ __forceinline__ void stg(GLOBAL_AS T* base, uint32_t boff, T v) {
    *(GLOBAL_AS T*)((GLOBAL_AS char*)base + boff) = v;
}
// The same store with the nontemporal hint (NT = true): streamed fields that will not be reread
// before they leave the caches.
template <bool NT, class T>
__device__ __forceinline__ void stf(GLOBAL_AS T* base, uint32_t boff, T v) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, (GLOBAL_AS T*)((GLOBAL_AS char*)base + boff));
    else
        stg(base, boff, v);
}

// A load with the nontemporal hint (NL = true): arrays streamed once per step (split fields, raw
// coefficient arrays, entry indices), whose lines need not stay in the caches.
template <bool NL, class T>
__device__ __forceinline__ T ldn(const GLOBAL_AS T* base, uint32_t boff) {
    if constexpr (NL)
        return __builtin_nontemporal_load((const GLOBAL_AS T*)((const GLOBAL_AS char*)base + boff));
    else
        return ldg(base, boff);
}

// E coefficients of one cell, order of the reference dictionary entries 0..11.
template <typename T>
struct ECoef {
    T ca_xy, ca_xz, cb_xy, cb_xz, ca_yz, ca_yx, cb_yz, cb_yx, ca_zx, ca_zy, cb_zx, cb_zy;
};
// H coefficients of one cell, entries 12..29.
template <typename T>
struct HCoef {
    T da_xy, da_xz, db_xy1, db_xy2, db_xz1, db_xz2;
    T da_yz, da_yx, db_yz1, db_yz2, db_yx1, db_yx2;
    T da_zx, da_zy, db_zx1, db_zx2, db_zy1, db_zy2;
};

// One instance (a grid and one excitation) as the kernels see it.
template <typename T>
struct InstDev {
    Fields<T> F;             // the fields the step reads (and the separate kernels update in place)
    Fields<T> G;             // fused step: the fields it writes (the other buffer of the pair)
    const T* raw[30];        // coefficient arrays (coef_kind 0)
    const uint32_t* eidx;    // per-cell E entry (coef_kind 1)
    const uint32_t* hidx;    // per-cell H entry
    const Vec4<T>* etab;     // 3 Vec4 per E entry
    const Vec4<T>* htab;     // 5 Vec4 per H entry (18 values + 2 pad)
    const uint32_t* ebits;   // one bit per cell holding an E probe
    const void* egroups;     // the E probe groups, sorted by cell
    const uint32_t* hbits;   // separate kernels: one bit per cell holding an H probe
    const void* hgroups;     // the H probe groups, sorted by cell
    int32_t coef_kind, Nx, Ny, Nz, xplanes, npb, tz_n, negroups, nhgroups, pad_;
};

// A member (grid) owns tiles * K logical workgroups starting at tile0; its K instances are
// consecutive from inst0 and vary fastest in the logical numbering.
struct MemberDev {
    int32_t tile0, tiles, K, inst0;
    int32_t order;  // the column kernels' workgroup numbering: 0 grouped by XCD, 1 launch order
};

// The 30 arrays as given (one coalesced load per coefficient).
template <typename T>
struct CoefRaw {
    const T* const* cp;
    // b: byte offset of the cell in every array
    __device__ __forceinline__ ECoef<T> e(uint32_t b) const {
        auto c = [&](int k) { return ldg((const GLOBAL_AS T*)cp[k], b); };
        ECoef<T> r;
        r.ca_xy = c(0); r.ca_xz = c(1); r.cb_xy = c(2); r.cb_xz = c(3);
        r.ca_yz = c(4); r.ca_yx = c(5); r.cb_yz = c(6); r.cb_yx = c(7);
        r.ca_zx = c(8); r.ca_zy = c(9); r.cb_zx = c(10); r.cb_zy = c(11);
        return r;
    }
    // coefficient k (0..29) alone
    __device__ __forceinline__ T at(uint32_t b, int k) const { return ldg((const GLOBAL_AS T*)cp[k], b); }
    // two-stage access (key first, values later) for software pipelining: here the key is the offset
    __device__ __forceinline__ uint32_t ekey(uint32_t b) const { return b; }
    __device__ __forceinline__ uint32_t hkey(uint32_t b) const { return b; }
    __device__ __forceinline__ T at_key(uint32_t key, int k) const { return at(key, k); }
    __device__ __forceinline__ ECoef<T> e_at(uint32_t key) const { return e(key); }
    __device__ __forceinline__ HCoef<T> h_at(uint32_t key) const { return h(key); }
    // the same loads with the nontemporal hint when NL
    template <bool NL>
    __device__ __forceinline__ ECoef<T> e_s(uint32_t b) const {
        auto c = [&](int k) { return ldn<NL>((const GLOBAL_AS T*)cp[k], b); };
        ECoef<T> r;
        r.ca_xy = c(0); r.ca_xz = c(1); r.cb_xy = c(2); r.cb_xz = c(3);
        r.ca_yz = c(4); r.ca_yx = c(5); r.cb_yz = c(6); r.cb_yx = c(7);
        r.ca_zx = c(8); r.ca_zy = c(9); r.cb_zx = c(10); r.cb_zy = c(11);
        return r;
    }
    template <bool NL>
    __device__ __forceinline__ HCoef<T> h_s(uint32_t b) const {
        auto c = [&](int k) { return ldn<NL>((const GLOBAL_AS T*)cp[k], b); };
        HCoef<T> r;
        r.da_xy = c(12); r.da_xz = c(13); r.db_xy1 = c(14); r.db_xy2 = c(15); r.db_xz1 = c(16); r.db_xz2 = c(17);
        r.da_yz = c(18); r.da_yx = c(19); r.db_yz1 = c(20); r.db_yz2 = c(21); r.db_yx1 = c(22); r.db_yx2 = c(23);
        r.da_zx = c(24); r.da_zy = c(25); r.db_zx1 = c(26); r.db_zx2 = c(27); r.db_zy1 = c(28); r.db_zy2 = c(29);
        return r;
    }
    __device__ __forceinline__ HCoef<T> h(uint32_t b) const {
        auto c = [&](int k) { return ldg((const GLOBAL_AS T*)cp[k], b); };
        HCoef<T> r;
        r.da_xy = c(12); r.da_xz = c(13); r.db_xy1 = c(14); r.db_xy2 = c(15); r.db_xz1 = c(16); r.db_xz2 = c(17);
        r.da_yz = c(18); r.da_yx = c(19); r.db_yz1 = c(20); r.db_yz2 = c(21); r.db_yx1 = c(22); r.db_yx2 = c(23);
        r.da_zx = c(24); r.da_zy = c(25); r.db_zx1 = c(26); r.db_zx2 = c(27); r.db_zy1 = c(28); r.db_zy2 = c(29);
        return r;
    }
};

// Deduplicated coefficients: per-cell entry index into tables of distinct tuples.
template <typename T>
struct CoefTable {
    const uint32_t* eidx;
    const uint32_t* hidx;
    const Vec4<T>* etab;
    const Vec4<T>* htab;
    // coefficient k (0..29) alone: E entries hold 12 values, H entries 20 (18 and padding)
    __device__ __forceinline__ T at(uint32_t bc, int k) const {
        const uint32_t S = (uint32_t)sizeof(T);
        if (k < 12) {
            const uint32_t e = ldg((const GLOBAL_AS uint32_t*)eidx, bc / S * 4u);
            return ldg((const GLOBAL_AS T*)etab, (e * 12u + (uint32_t)k) * S);
        }
        const uint32_t e = ldg((const GLOBAL_AS uint32_t*)hidx, bc / S * 4u);
        return ldg((const GLOBAL_AS T*)htab, (e * 20u + (uint32_t)(k - 12)) * S);
    }
    // two-stage access (entry index first, values later) for software pipelining
    __device__ __forceinline__ uint32_t ekey(uint32_t bc) const {
        return ldg((const GLOBAL_AS uint32_t*)eidx, bc / (uint32_t)sizeof(T) * 4u);
    }
    __device__ __forceinline__ uint32_t hkey(uint32_t bc) const {
        return ldg((const GLOBAL_AS uint32_t*)hidx, bc / (uint32_t)sizeof(T) * 4u);
    }
    // coefficient k alone from an entry index (an E entry for k < 12, else an H entry)
    __device__ __forceinline__ T at_key(uint32_t key, int k) const {
        if (k < 12) return ldg((const GLOBAL_AS T*)etab, (key * 12u + (uint32_t)k) * (uint32_t)sizeof(T));
        return ldg((const GLOBAL_AS T*)htab, (key * 20u + (uint32_t)(k - 12)) * (uint32_t)sizeof(T));
    }
    __device__ __forceinline__ ECoef<T> e(uint32_t bc) const { return e_at(ekey(bc)); }
    __device__ __forceinline__ HCoef<T> h(uint32_t bc) const { return h_at(hkey(bc)); }
    // the entry index streamed (nontemporal hint when NL), the table entry cached
    template <bool NL>
    __device__ __forceinline__ ECoef<T> e_s(uint32_t bc) const {
        return e_at(ldn<NL>((const GLOBAL_AS uint32_t*)eidx, bc / (uint32_t)sizeof(T) * 4u));
    }
    template <bool NL>
    __device__ __forceinline__ HCoef<T> h_s(uint32_t bc) const {
        return h_at(ldn<NL>((const GLOBAL_AS uint32_t*)hidx, bc / (uint32_t)sizeof(T) * 4u));
    }
    __device__ __forceinline__ ECoef<T> e_at(uint32_t k) const {
        const GLOBAL_AS Vec4<T>* tb = (const GLOBAL_AS Vec4<T>*)etab;
        const uint32_t o = k * (uint32_t)(3 * sizeof(Vec4<T>));
        const Vec4<T> a = ldg(tb, o), b = ldg(tb, o + (uint32_t)sizeof(Vec4<T>)),
                      c = ldg(tb, o + 2u * (uint32_t)sizeof(Vec4<T>));
        ECoef<T> r;
        r.ca_xy = a.v[0]; r.ca_xz = a.v[1]; r.cb_xy = a.v[2]; r.cb_xz = a.v[3];
        r.ca_yz = b.v[0]; r.ca_yx = b.v[1]; r.cb_yz = b.v[2]; r.cb_yx = b.v[3];
        r.ca_zx = c.v[0]; r.ca_zy = c.v[1]; r.cb_zx = c.v[2]; r.cb_zy = c.v[3];
        return r;
    }
    __device__ __forceinline__ HCoef<T> h_at(uint32_t k) const {
        const GLOBAL_AS Vec4<T>* tb = (const GLOBAL_AS Vec4<T>*)htab;
        const uint32_t o = k * (uint32_t)(5 * sizeof(Vec4<T>)), v = (uint32_t)sizeof(Vec4<T>);
        const Vec4<T> a = ldg(tb, o), b = ldg(tb, o + v), c = ldg(tb, o + 2u * v), d = ldg(tb, o + 3u * v),
                      e = ldg(tb, o + 4u * v);
        HCoef<T> r;
        r.da_xy = a.v[0]; r.da_xz = a.v[1]; r.db_xy1 = a.v[2]; r.db_xy2 = a.v[3];
        r.db_xz1 = b.v[0]; r.db_xz2 = b.v[1];
        r.da_yz = b.v[2]; r.da_yx = b.v[3]; r.db_yz1 = c.v[0]; r.db_yz2 = c.v[1];
        r.db_yx1 = c.v[2]; r.db_yx2 = c.v[3];
        r.da_zx = d.v[0]; r.da_zy = d.v[1]; r.db_zx1 = d.v[2]; r.db_zx2 = d.v[3];
        r.db_zy1 = e.v[0]; r.db_zy2 = e.v[1];
        return r;
    }
};

// Workgroups are dispatched round-robin over the 8 XCDs (id -> XCD id % 8).  Renumber them so that
// each XCD works through one contiguous range of logical tiles; the launcher pads the grid to a
// multiple of 8 and surplus ids exit.
__device__ __forceinline__ int xcd_logical(int bid, int nblocks) {
    const int per = nblocks >> 3;
    return (bid & 7) * per + (bid >> 3);
}

// The column kernels (the separate E and H updates, their vector forms, the adjoint passes) number
// their workgroups either way, as the launch's member table says (the same for every member):
// grouped by XCD, or in launch order, where consecutive workgroups, which the hardware spreads over
// the XCDs, stream neighbouring parts of the arrays together.  On a 32 M-cell grid launch order is
// 3-10% faster (a plain streaming kernel too, by 3.5%); on grids of a few M cells and less the XCD
// grouping is, as the y-neighbour rows its tiles share stay in one L2.  The tuner picks per session.
// The tiled passes always group by XCD: their neighbouring tiles share halo rows through the L2.
__device__ __forceinline__ int column_block(const MemberDev* __restrict__ mem) {
    return mem[0].order ? (int)blockIdx.x : xcd_logical(blockIdx.x, gridDim.x);
}

struct Tile {
    int64_t p;  // in-plane offset y*Nz + z
    int y, z, x0, x1;
};

// Map this workgroup to (instance, tile) and this thread to its (y,z) column and x planes.
// Returns nullptr for threads with no cell.  The member lookup is uniform per workgroup.
template <typename T>
__device__ __forceinline__ const InstDev<T>* locate(const InstDev<T>* __restrict__ insts,
                                                     const MemberDev* __restrict__ mem, int nmem, int total,
                                                     int* s_member, Tile& t, int W = 1) {
    const int L = column_block(mem);
    if (threadIdx.x == 0) {
        int m = 0;
        if (L < total)
            while (m + 1 < nmem && mem[m + 1].tile0 <= L) ++m;
        *s_member = m;
    }
    __syncthreads();
    if (L >= total) return nullptr;
    // the member index comes through LDS: tell the compiler it is wave-uniform, so the member and
    // instance descriptors are fetched with scalar loads into SGPRs
    const int m = __builtin_amdgcn_readfirstlane(*s_member);
    const MemberDev M = mem[m];
    const int local = L - M.tile0;
    const int e = local % M.K, tile = local / M.K;
    const InstDev<T>* I = insts + __builtin_amdgcn_readfirstlane(M.inst0 + e);
    const int Ny = I->Ny, Nz = I->Nz, npb = I->npb;
    const int pb = tile % npb, xc = tile / npb;
    t.p = ((int64_t)pb * blockDim.x + threadIdx.x) * W;  // W consecutive cells of a row per thread
    t.y = (int)(t.p / Nz);
    t.z = (int)(t.p - (int64_t)t.y * Nz);
    t.x0 = xc * I->xplanes;
    // lanes past the last (y,z) column get an empty x range: the pointer stays wave-uniform
    t.x1 = t.p < (int64_t)Ny * Nz ? min(I->Nx, t.x0 + I->xplanes) : t.x0;
    return I;
}

// Field and coefficient pointers reach the kernels through descriptors in memory, so the compiler
// sees generic pointers; casting them to the global address space gives global (not flat) loads
// with a scalar base and a 32-bit vector offset.
#define GLOBAL_T(T) GLOBAL_AS T

// Probes sharing one (component, cell) of one instance form a group processed by one thread in
// list order.  series is the instance's time-major [Nt][nprobes] array.
template <typename T>
struct ProbeGroup {
    int64_t cell;
    T* series;              // in/out: source samples in, sampled totals out
    const T* series_in;     // pristine copy of the source samples (recomputation)
    const int32_t* order;   // instance-local probe ids of the group, in list order
    const int32_t* is_src;  // per instance-local probe id
    int32_t inst, field, count, nprobes;
};

__device__ __forceinline__ int step_of(const int* base, int local) { return base ? (*base + local) : local; }

// Sources and records of the probe groups at one cell (groups sorted by cell, stably, so groups
// keep list order) applied to the six split values of E, or of H, just computed there, with the
// source samples from the pristine series: per probe in list order, a source adds its sample to
// both splits (s1 += v, s2 += v, total = s1 + s2) and every probe records the total
// (rfnetwork/core/src/solver.cpp:963-976, :1104-1116).
template <typename T>
__device__ __forceinline__ void inject_cell(const ProbeGroup<T>* __restrict__ g, int ng, int64_t cell, int n,
                                            bool record, T& xy, T& xz, T& yz, T& yx, T& zx, T& zy) {
    int lo = 0, hi = ng;
    while (lo < hi) {  // first group at this cell
        const int mid = (lo + hi) >> 1;
        if (g[mid].cell < cell) lo = mid + 1;
        else hi = mid;
    }
    for (int k = lo; k < ng && g[k].cell == cell; ++k) {
        const ProbeGroup<T> G = g[k];
        const int c = G.field % 3;  // x, y or z component of E or H
        T& a = c == 0 ? xy : c == 1 ? yz : zx;
        T& b = c == 0 ? xz : c == 1 ? yx : zy;
        T s = a + b;
        T* row = G.series + (size_t)n * G.nprobes;
        const T* row_in = G.series_in + (size_t)n * G.nprobes;
        for (int j = 0; j < G.count; ++j) {
            const int pid = G.order[j];
            if (G.is_src[pid]) {
                const T v = row_in[pid];
                a = a + v;
                b = b + v;
                s = a + b;
            }
            if (record) row[pid] = s;
        }
    }
}

// The probes of one field (E or H) folded into its update kernel: a bit per cell marks probe
// cells, whose groups (sorted by cell) are applied to the values before they are stored.
// ngroups 0: nothing folded (tuning, recomputation, or no probes).
template <typename T>
struct ProbeFold {
    const uint32_t* bits;
    const ProbeGroup<T>* groups;
    int ngroups, n;
    // the bitmask word of a cell, loaded with the rest of its plane's inputs (0 when nothing is folded)
    __device__ __forceinline__ uint32_t word(uint32_t cell) const {
        return ngroups ? ldg((const GLOBAL_AS uint32_t*)bits, (cell >> 5) * 4u) : 0u;
    }
    __device__ __forceinline__ void apply(uint32_t cell, uint32_t w, T& xy, T& xz, T& yz, T& yx, T& zx, T& zy) const {
        if ((w >> (cell & 31)) & 1u) inject_cell(groups, ngroups, (int64_t)cell, n, true, xy, xz, yz, yx, zx, zy);
    }
};

// Adjoint sources folded into the reverse kernels: for every group at a marked cell, g[n] summed
// over the group's probes (series holds dJ/dy, time-major) and added to both splits of its
// component, as k_adj_inject does.  ngroups 0: nothing folded.
template <typename T>
struct AdjFold {
    const uint32_t* bits;
    const ProbeGroup<T>* groups;
    int ngroups, n;
    __device__ __forceinline__ uint32_t word(uint32_t cell) const {
        return ngroups ? ldg((const GLOBAL_AS uint32_t*)bits, (cell >> 5) * 4u) : 0u;
    }
    __device__ __forceinline__ void apply(uint32_t cell, T& xy, T& xz, T& yz, T& yx, T& zx, T& zy) const {
        apply(cell, word(cell), xy, xz, yz, yx, zx, zy);
    }
    // w: the cell's bitmask word, loaded ahead with the plane's inputs
    __device__ __forceinline__ void apply(uint32_t cell, uint32_t w, T& xy, T& xz, T& yz, T& yx, T& zx, T& zy) const {
        if (!((w >> (cell & 31)) & 1u)) return;
        int lo = 0, hi = ngroups;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (groups[mid].cell < (int64_t)cell) lo = mid + 1;
            else hi = mid;
        }
        for (int k = lo; k < ngroups && groups[k].cell == (int64_t)cell; ++k) {
            const ProbeGroup<T> G = groups[k];
            const T* row = G.series + (size_t)n * G.nprobes;
            T sum = T(0);
            for (int j = 0; j < G.count; ++j) sum = sum + row[G.order[j]];
            const int c = G.field % 3;
            T& a = c == 0 ? xy : c == 1 ? yz : zx;
            T& b = c == 0 ? xz : c == 1 ? yx : zy;
            a = a + sum;
            b = b + sum;
        }
    }
};

struct MonitorDev {
    int32_t inst, axis, pos, field, n_step, nf, N1, N2, blk0, id, fblk0, ring;  // ring: samples kept (0: all)
    int32_t acc64;  // running DFT accumulated in double (values and phase double2) in a float session
    int64_t N1N2;
    void* values;
    const void* phase;
    void* stage;  // phasor monitors: kStageSlots plane samples awaiting accumulation
};

// Running DFTs are accumulated every kStageSlots steps instead of every step: each step copies the
// monitor plane into a staging slot, and a flush adds the staged samples to the accumulators in time
// order, so the sums are the same float operations in the same order while the accumulators (Nf
// complex values per plane cell) are read and written once per kStageSlots steps.
constexpr int kStageSlots = 20;
constexpr int kFreqGroup = 8;  // frequencies whose accumulators the flush loads together

// Everything one x plane of the E update reads, so that the next plane's loads can be issued
// before this plane's arithmetic and stores (PF = true: two planes of loads in flight per lane).
template <typename T>
struct EIn {
    T hx_c, hy_xn, hz_xn, hz_yn, hx_yn, hy_zn, hx_zn;
    T exy, exz, eyz, eyx, ezx, ezy;
    ECoef<T> c;
    uint32_t pw;  // probe bitmask word of the cell
};

template <typename T>
struct HIn {
    T ex_c, ey_c, ez_c, ez_ym, ex_ym, ey_zm, ex_zm;
    T hxy, hxz, hyz, hyx, hzx, hzy;
    HCoef<T> c;
    uint32_t pw;  // probe bitmask word of the cell
};

// E half step over one column of x planes.  Offsets are 32-bit byte offsets (grids up to 2^29
// cells).
template <typename T, class CA, bool PF, bool NT, bool NL = false>
__device__ __forceinline__ void e_column(const Fields<T>& F, const CA& C, int Nx, int Ny, int Nz, const Tile& t,
                                         const ProbeFold<T>& P) {
    using G = GLOBAL_T(T);
    G* const ex_y = (G*)F.ex_y; G* const ex_z = (G*)F.ex_z;
    G* const ey_z = (G*)F.ey_z; G* const ey_x = (G*)F.ey_x;
    G* const ez_x = (G*)F.ez_x; G* const ez_y = (G*)F.ez_y;
    G* const ex = (G*)F.ex; G* const ey = (G*)F.ey; G* const ez = (G*)F.ez;
    const G* const hx = (const G*)F.hx; const G* const hy = (const G*)F.hy; const G* const hz = (const G*)F.hz;
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S, sy = (uint32_t)Nz * S;
    const bool yin = t.y < Ny - 1, zin = t.z < Nz - 1;
    auto load = [&](EIn<T>& v, uint32_t b, bool xin) {
        v.hx_c = ldg(hx, b);
        v.hy_xn = xin ? ldg(hy, b + sx) : T(0);
        v.hz_xn = xin ? ldg(hz, b + sx) : T(0);
        v.hz_yn = yin ? ldg(hz, b + sy) : T(0);
        v.hx_yn = yin ? ldg(hx, b + sy) : T(0);
        v.hy_zn = zin ? ldg(hy, b + S) : T(0);
        v.hx_zn = zin ? ldg(hx, b + S) : T(0);
        v.c = C.template e_s<NL>(b);
        v.pw = P.word(b / S);
        v.exy = ldn<NL>(ex_y, b); v.exz = ldn<NL>(ex_z, b); v.eyz = ldn<NL>(ey_z, b);
        v.eyx = ldn<NL>(ey_x, b); v.ezx = ldn<NL>(ez_x, b); v.ezy = ldn<NL>(ez_y, b);
    };
    if (t.x0 >= t.x1) return;
    uint32_t b = ((uint32_t)t.x0 * (uint32_t)Ny * (uint32_t)Nz + (uint32_t)t.p) * S;
    T hy_c = ldg(hy, b), hz_c = ldg(hz, b);
    EIn<T> cur;
    load(cur, b, t.x0 < Nx - 1);
    for (int x = t.x0; x < t.x1; ++x, b += sx) {
        const bool xin = x < Nx - 1, more = x + 1 < t.x1;
        EIn<T> nxt;
        if (PF && more) load(nxt, b + sx, x + 1 < Nx - 1);
        const ECoef<T>& c = cur.c;
        T exy = cur.exy, exz = cur.exz, eyz = cur.eyz, eyx = cur.eyx, ezx = cur.ezx, ezy = cur.ezy;
        if (yin) exy = c.ca_xy * exy + c.cb_xy * (cur.hz_yn - hz_c);
        if (zin) exz = c.ca_xz * exz + c.cb_xz * (cur.hy_zn - hy_c);
        if (zin) eyz = c.ca_yz * eyz + c.cb_yz * (cur.hx_zn - cur.hx_c);
        if (xin) eyx = c.ca_yx * eyx + c.cb_yx * (cur.hz_xn - hz_c);
        if (xin) ezx = c.ca_zx * ezx + c.cb_zx * (cur.hy_xn - hy_c);
        if (yin) ezy = c.ca_zy * ezy + c.cb_zy * (cur.hx_yn - cur.hx_c);
        P.apply(b / S, cur.pw, exy, exz, eyz, eyx, ezx, ezy);
        stf<NT>(ex_y, b, exy); stf<NT>(ex_z, b, exz); stf<NT>(ey_z, b, eyz);
        stf<NT>(ey_x, b, eyx); stf<NT>(ez_x, b, ezx); stf<NT>(ez_y, b, ezy);
        stf<NT>(ex, b, exy + exz);
        stf<NT>(ey, b, eyz + eyx);
        stf<NT>(ez, b, ezx + ezy);
        hy_c = cur.hy_xn;
        hz_c = cur.hz_xn;
        if (PF)
            cur = nxt;
        else if (more)
            load(cur, b + sx, x + 1 < Nx - 1);
    }
}

// H half step over one column of x planes; neighbours below index 0 are zero.
template <typename T, class CA, bool PF, bool NT, bool NL = false>
__device__ __forceinline__ void h_column(const Fields<T>& F, const CA& C, int Nx, int Ny, int Nz, const Tile& t,
                                         const ProbeFold<T>& P) {
    using G = GLOBAL_T(T);
    G* const hx_y = (G*)F.hx_y; G* const hx_z = (G*)F.hx_z;
    G* const hy_z = (G*)F.hy_z; G* const hy_x = (G*)F.hy_x;
    G* const hz_x = (G*)F.hz_x; G* const hz_y = (G*)F.hz_y;
    G* const hx = (G*)F.hx; G* const hy = (G*)F.hy; G* const hz = (G*)F.hz;
    const G* const ex = (const G*)F.ex; const G* const ey = (const G*)F.ey; const G* const ez = (const G*)F.ez;
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S, sy = (uint32_t)Nz * S;
    const bool ylo = t.y > 0, zlo = t.z > 0;
    (void)Nx;
    auto load = [&](HIn<T>& v, uint32_t b) {
        v.ex_c = ldg(ex, b); v.ey_c = ldg(ey, b); v.ez_c = ldg(ez, b);
        v.ez_ym = ylo ? ldg(ez, b - sy) : T(0);
        v.ex_ym = ylo ? ldg(ex, b - sy) : T(0);
        v.ey_zm = zlo ? ldg(ey, b - S) : T(0);
        v.ex_zm = zlo ? ldg(ex, b - S) : T(0);
        v.c = C.template h_s<NL>(b);
        v.pw = P.word(b / S);
        v.hxy = ldn<NL>(hx_y, b); v.hxz = ldn<NL>(hx_z, b); v.hyz = ldn<NL>(hy_z, b);
        v.hyx = ldn<NL>(hy_x, b); v.hzx = ldn<NL>(hz_x, b); v.hzy = ldn<NL>(hz_y, b);
    };
    if (t.x0 >= t.x1) return;
    uint32_t b = ((uint32_t)t.x0 * (uint32_t)Ny * (uint32_t)Nz + (uint32_t)t.p) * S;
    T ey_xm = t.x0 > 0 ? ldg(ey, b - sx) : T(0);
    T ez_xm = t.x0 > 0 ? ldg(ez, b - sx) : T(0);
    HIn<T> cur;
    load(cur, b);
    for (int x = t.x0; x < t.x1; ++x, b += sx) {
        const bool more = x + 1 < t.x1;
        HIn<T> nxt;
        if (PF && more) load(nxt, b + sx);
        const HCoef<T>& c = cur.c;
        T hxy = c.da_xy * cur.hxy + (c.db_xy2 * cur.ez_c - c.db_xy1 * cur.ez_ym);
        T hxz = c.da_xz * cur.hxz + (c.db_xz2 * cur.ey_c - c.db_xz1 * cur.ey_zm);
        T hyz = c.da_yz * cur.hyz + (c.db_yz2 * cur.ex_c - c.db_yz1 * cur.ex_zm);
        T hyx = c.da_yx * cur.hyx + (c.db_yx2 * cur.ez_c - c.db_yx1 * ez_xm);
        T hzx = c.da_zx * cur.hzx + (c.db_zx2 * cur.ey_c - c.db_zx1 * ey_xm);
        T hzy = c.da_zy * cur.hzy + (c.db_zy2 * cur.ex_c - c.db_zy1 * cur.ex_ym);
        P.apply(b / S, cur.pw, hxy, hxz, hyz, hyx, hzx, hzy);
        stf<NT>(hx_y, b, hxy); stf<NT>(hx_z, b, hxz); stf<NT>(hy_z, b, hyz);
        stf<NT>(hy_x, b, hyx); stf<NT>(hz_x, b, hzx); stf<NT>(hz_y, b, hzy);
        stf<NT>(hx, b, hxy + hxz);
        stf<NT>(hy, b, hyz + hyx);
        stf<NT>(hz, b, hzx + hzy);
        ey_xm = cur.ey_c;
        ez_xm = cur.ez_c;
        if (PF)
            cur = nxt;
        else if (more)
            load(cur, b + sx);
    }
}

// io != 0: the instance's E (H) probes are applied inside the kernel at step base/local.
template <typename T, bool PF, bool NT = false, bool NL = false>
__global__ void __launch_bounds__(1024) k_update_e(const InstDev<T>* __restrict__ insts,
                                                   const MemberDev* __restrict__ mem, int nmem, int total,
                                                   const int* base, int local, int io) {
    __shared__ int s_member;
    Tile t;
    const InstDev<T>* I = locate(insts, mem, nmem, total, &s_member, t);
    if (!I) return;
    const int ng = io ? I->negroups : 0;
    const ProbeFold<T> P{I->ebits, (const ProbeGroup<T>*)I->egroups, ng, ng ? step_of(base, local) : 0};
    if (I->coef_kind == 1)
        e_column<T, CoefTable<T>, PF, NT, NL>(I->F, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, I->Nx, I->Ny,
                                              I->Nz, t, P);
    else
        e_column<T, CoefRaw<T>, PF, NT, NL>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
}

template <typename T, bool PF, bool NT = false, bool NL = false>
__global__ void __launch_bounds__(1024) k_update_h(const InstDev<T>* __restrict__ insts,
                                                   const MemberDev* __restrict__ mem, int nmem, int total,
                                                   const int* base, int local, int io) {
    __shared__ int s_member;
    Tile t;
    const InstDev<T>* I = locate(insts, mem, nmem, total, &s_member, t);
    if (!I) return;
    const int ng = io ? I->nhgroups : 0;
    const ProbeFold<T> P{I->hbits, (const ProbeGroup<T>*)I->hgroups, ng, ng ? step_of(base, local) : 0};
    if (I->coef_kind == 1)
        h_column<T, CoefTable<T>, PF, NT, NL>(I->F, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, I->Nx, I->Ny,
                                              I->Nz, t, P);
    else
        h_column<T, CoefRaw<T>, PF, NT, NL>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
}


// ------------------------------------------------------------------------------------------------
// separate E and H kernels over W consecutive z cells per thread (variants 11 and 12, W = 2 and 4)
//
// The same columns, arithmetic and order as e_column / h_column, with every array streamed once per
// step read and written as one W-wide vector per thread (global_load_dwordx2 / x4): a wavefront
// moves W x 256 bytes per load instruction, so each lane keeps W times the bytes in flight for the
// same number of instructions.  The z neighbours inside a group come from the group's own vector;
// the one past its end (E: z + W) or before its start (H: z - 1) is one more scalar load, an L1 hit
// since the neighbouring lane loads that line.  Requires Nz % W == 0, so that a group stays in one
// row and every vector is aligned; the groups' probe cells share one 32-cell bitmask word.
// ------------------------------------------------------------------------------------------------

template <typename T, int W>
struct VecOf {
    typedef T type __attribute__((ext_vector_type(W)));
};

// W consecutive values at a byte offset from a wave-uniform base (nontemporal hint when NL)
template <bool NL, int W, class T>
__device__ __forceinline__ void ldw(const GLOBAL_AS T* base, uint32_t boff, T (&out)[W]) {
    using V = typename VecOf<T, W>::type;
    const GLOBAL_AS V* q = (const GLOBAL_AS V*)((const GLOBAL_AS char*)base + boff);
    V v;
    if constexpr (NL)
        v = __builtin_nontemporal_load(q);
    else
        v = *q;
#pragma unroll
    for (int k = 0; k < W; ++k) out[k] = v[k];
}
template <int W, class T>
__device__ __forceinline__ void stw(GLOBAL_AS T* base, uint32_t boff, const T (&in)[W]) {
    using V = typename VecOf<T, W>::type;
    V v;
#pragma unroll
    for (int k = 0; k < W; ++k) v[k] = in[k];
    __builtin_nontemporal_store(v, (GLOBAL_AS V*)((GLOBAL_AS char*)base + boff));
}

// The 12 E (18 H) coefficients of the W cells of a group: coefficient k of cell i in c[k][i].
template <typename T, int W>
__device__ __forceinline__ void coef_e_w(const CoefRaw<T>& C, uint32_t b, T (&c)[12][W]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) ldw<true, W>((const GLOBAL_AS T*)C.cp[k], b, c[k]);
}
template <typename T, int W>
__device__ __forceinline__ void coef_e_w(const CoefTable<T>& C, uint32_t b, T (&c)[12][W]) {
    uint32_t key[W];
    ldw<true, W>((const GLOBAL_AS uint32_t*)C.eidx, b / (uint32_t)sizeof(T) * 4u, key);
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const ECoef<T> e = C.e_at(key[i]);
        c[0][i] = e.ca_xy; c[1][i] = e.ca_xz; c[2][i] = e.cb_xy; c[3][i] = e.cb_xz;
        c[4][i] = e.ca_yz; c[5][i] = e.ca_yx; c[6][i] = e.cb_yz; c[7][i] = e.cb_yx;
        c[8][i] = e.ca_zx; c[9][i] = e.ca_zy; c[10][i] = e.cb_zx; c[11][i] = e.cb_zy;
    }
}
template <typename T, int W>
__device__ __forceinline__ void coef_h_w(const CoefRaw<T>& C, uint32_t b, T (&c)[18][W]) {
#pragma unroll
    for (int k = 0; k < 18; ++k) ldw<true, W>((const GLOBAL_AS T*)C.cp[12 + k], b, c[k]);
}
template <typename T, int W>
__device__ __forceinline__ void coef_h_w(const CoefTable<T>& C, uint32_t b, T (&c)[18][W]) {
    uint32_t key[W];
    ldw<true, W>((const GLOBAL_AS uint32_t*)C.hidx, b / (uint32_t)sizeof(T) * 4u, key);
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const HCoef<T> h = C.h_at(key[i]);
        c[0][i] = h.da_xy; c[1][i] = h.da_xz; c[2][i] = h.db_xy1; c[3][i] = h.db_xy2; c[4][i] = h.db_xz1;
        c[5][i] = h.db_xz2; c[6][i] = h.da_yz; c[7][i] = h.da_yx; c[8][i] = h.db_yz1; c[9][i] = h.db_yz2;
        c[10][i] = h.db_yx1; c[11][i] = h.db_yx2; c[12][i] = h.da_zx; c[13][i] = h.da_zy; c[14][i] = h.db_zx1;
        c[15][i] = h.db_zx2; c[16][i] = h.db_zy1; c[17][i] = h.db_zy2;
    }
}

template <typename T, int W>
struct EInW {
    T hx_c[W], hy_xn[W], hz_xn[W], hz_yn[W], hx_yn[W];
    T hy_zw, hx_zw;  // hy, hx at z + W (the next group's first cell)
    T e[6][W];       // ex_y, ex_z, ey_z, ey_x, ez_x, ez_y
    T c[12][W];
    uint32_t pw;
};

template <typename T, int W>
struct HInW {
    T ex_c[W], ey_c[W], ez_c[W], ez_ym[W], ex_ym[W];
    T ey_z1, ex_z1;  // ey, ex at z - 1 (the previous group's last cell)
    T h[6][W];       // hx_y, hx_z, hy_z, hy_x, hz_x, hz_y
    T c[18][W];
    uint32_t pw;
};

template <typename T, class CA, int W, bool PF>
__device__ __forceinline__ void e_column_w(const Fields<T>& F, const CA& C, int Nx, int Ny, int Nz, const Tile& t,
                                           const ProbeFold<T>& P) {
    using G = GLOBAL_T(T);
    G* const es[6] = {(G*)F.ex_y, (G*)F.ex_z, (G*)F.ey_z, (G*)F.ey_x, (G*)F.ez_x, (G*)F.ez_y};
    G* const ex = (G*)F.ex; G* const ey = (G*)F.ey; G* const ez = (G*)F.ez;
    const G* const hx = (const G*)F.hx; const G* const hy = (const G*)F.hy; const G* const hz = (const G*)F.hz;
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S, sy = (uint32_t)Nz * S;
    const bool yin = t.y < Ny - 1;
    const bool next = t.z + W < Nz;  // the next group is in this row
    auto load = [&](EInW<T, W>& v, uint32_t b, bool xin) {
        ldw<false, W>(hx, b, v.hx_c);
        if (xin) {
            ldw<false, W>(hy, b + sx, v.hy_xn);
            ldw<false, W>(hz, b + sx, v.hz_xn);
        } else {
#pragma unroll
            for (int i = 0; i < W; ++i) v.hy_xn[i] = v.hz_xn[i] = T(0);
        }
        if (yin) {
            ldw<false, W>(hz, b + sy, v.hz_yn);
            ldw<false, W>(hx, b + sy, v.hx_yn);
        } else {
#pragma unroll
            for (int i = 0; i < W; ++i) v.hz_yn[i] = v.hx_yn[i] = T(0);
        }
        v.hy_zw = next ? ldg(hy, b + W * S) : T(0);
        v.hx_zw = next ? ldg(hx, b + W * S) : T(0);
        coef_e_w<T, W>(C, b, v.c);
        v.pw = P.word(b / S);
#pragma unroll
        for (int k = 0; k < 6; ++k) ldw<true, W>(es[k], b, v.e[k]);
    };
    if (t.x0 >= t.x1) return;
    uint32_t b = ((uint32_t)t.x0 * (uint32_t)Ny * (uint32_t)Nz + (uint32_t)t.p) * S;
    T hy_c[W], hz_c[W];
    ldw<false, W>(hy, b, hy_c);
    ldw<false, W>(hz, b, hz_c);
    EInW<T, W> cur;
    load(cur, b, t.x0 < Nx - 1);
    for (int x = t.x0; x < t.x1; ++x, b += sx) {
        const bool xin = x < Nx - 1, more = x + 1 < t.x1;
        EInW<T, W> nxt;
        if (PF && more) load(nxt, b + sx, x + 1 < Nx - 1);
        T o[6][W], tx[W], ty[W], tz[W];
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const bool zin = t.z + i < Nz - 1;
            const T hy_zn = i + 1 < W ? hy_c[i + 1] : cur.hy_zw;
            const T hx_zn = i + 1 < W ? cur.hx_c[i + 1] : cur.hx_zw;
            const T(&c)[12][W] = cur.c;
            T exy = cur.e[0][i], exz = cur.e[1][i], eyz = cur.e[2][i], eyx = cur.e[3][i], ezx = cur.e[4][i],
              ezy = cur.e[5][i];
            if (yin) exy = c[0][i] * exy + c[2][i] * (cur.hz_yn[i] - hz_c[i]);
            if (zin) exz = c[1][i] * exz + c[3][i] * (hy_zn - hy_c[i]);
            if (zin) eyz = c[4][i] * eyz + c[6][i] * (hx_zn - cur.hx_c[i]);
            if (xin) eyx = c[5][i] * eyx + c[7][i] * (cur.hz_xn[i] - hz_c[i]);
            if (xin) ezx = c[8][i] * ezx + c[10][i] * (cur.hy_xn[i] - hy_c[i]);
            if (yin) ezy = c[9][i] * ezy + c[11][i] * (cur.hx_yn[i] - cur.hx_c[i]);
            P.apply(b / S + i, cur.pw, exy, exz, eyz, eyx, ezx, ezy);
            o[0][i] = exy; o[1][i] = exz; o[2][i] = eyz; o[3][i] = eyx; o[4][i] = ezx; o[5][i] = ezy;
            tx[i] = exy + exz;
            ty[i] = eyz + eyx;
            tz[i] = ezx + ezy;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) stw<W>(es[k], b, o[k]);
        stw<W>(ex, b, tx);
        stw<W>(ey, b, ty);
        stw<W>(ez, b, tz);
#pragma unroll
        for (int i = 0; i < W; ++i) {
            hy_c[i] = cur.hy_xn[i];
            hz_c[i] = cur.hz_xn[i];
        }
        if (PF)
            cur = nxt;
        else if (more)
            load(cur, b + sx, x + 1 < Nx - 1);
    }
}

template <typename T, class CA, int W, bool PF>
__device__ __forceinline__ void h_column_w(const Fields<T>& F, const CA& C, int Nx, int Ny, int Nz, const Tile& t,
                                           const ProbeFold<T>& P) {
    using G = GLOBAL_T(T);
    G* const hs[6] = {(G*)F.hx_y, (G*)F.hx_z, (G*)F.hy_z, (G*)F.hy_x, (G*)F.hz_x, (G*)F.hz_y};
    G* const hx = (G*)F.hx; G* const hy = (G*)F.hy; G* const hz = (G*)F.hz;
    const G* const ex = (const G*)F.ex; const G* const ey = (const G*)F.ey; const G* const ez = (const G*)F.ez;
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S, sy = (uint32_t)Nz * S;
    const bool ylo = t.y > 0, zlo = t.z > 0;  // zlo: the previous group is in this row
    (void)Nx;
    auto load = [&](HInW<T, W>& v, uint32_t b) {
        ldw<false, W>(ex, b, v.ex_c);
        ldw<false, W>(ey, b, v.ey_c);
        ldw<false, W>(ez, b, v.ez_c);
        if (ylo) {
            ldw<false, W>(ez, b - sy, v.ez_ym);
            ldw<false, W>(ex, b - sy, v.ex_ym);
        } else {
#pragma unroll
            for (int i = 0; i < W; ++i) v.ez_ym[i] = v.ex_ym[i] = T(0);
        }
        v.ey_z1 = zlo ? ldg(ey, b - S) : T(0);
        v.ex_z1 = zlo ? ldg(ex, b - S) : T(0);
        coef_h_w<T, W>(C, b, v.c);
        v.pw = P.word(b / S);
#pragma unroll
        for (int k = 0; k < 6; ++k) ldw<true, W>(hs[k], b, v.h[k]);
    };
    if (t.x0 >= t.x1) return;
    uint32_t b = ((uint32_t)t.x0 * (uint32_t)Ny * (uint32_t)Nz + (uint32_t)t.p) * S;
    T ey_xm[W], ez_xm[W];
    if (t.x0 > 0) {
        ldw<false, W>(ey, b - sx, ey_xm);
        ldw<false, W>(ez, b - sx, ez_xm);
    } else {
#pragma unroll
        for (int i = 0; i < W; ++i) ey_xm[i] = ez_xm[i] = T(0);
    }
    HInW<T, W> cur;
    load(cur, b);
    for (int x = t.x0; x < t.x1; ++x, b += sx) {
        const bool more = x + 1 < t.x1;
        HInW<T, W> nxt;
        if (PF && more) load(nxt, b + sx);
        T o[6][W], tx[W], ty[W], tz[W];
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const T ey_zm = i > 0 ? cur.ey_c[i - 1] : cur.ey_z1;
            const T ex_zm = i > 0 ? cur.ex_c[i - 1] : cur.ex_z1;
            const T(&c)[18][W] = cur.c;
            T hxy = c[0][i] * cur.h[0][i] + (c[3][i] * cur.ez_c[i] - c[2][i] * cur.ez_ym[i]);
            T hxz = c[1][i] * cur.h[1][i] + (c[5][i] * cur.ey_c[i] - c[4][i] * ey_zm);
            T hyz = c[6][i] * cur.h[2][i] + (c[9][i] * cur.ex_c[i] - c[8][i] * ex_zm);
            T hyx = c[7][i] * cur.h[3][i] + (c[11][i] * cur.ez_c[i] - c[10][i] * ez_xm[i]);
            T hzx = c[12][i] * cur.h[4][i] + (c[15][i] * cur.ey_c[i] - c[14][i] * ey_xm[i]);
            T hzy = c[13][i] * cur.h[5][i] + (c[17][i] * cur.ex_c[i] - c[16][i] * cur.ex_ym[i]);
            P.apply(b / S + i, cur.pw, hxy, hxz, hyz, hyx, hzx, hzy);
            o[0][i] = hxy; o[1][i] = hxz; o[2][i] = hyz; o[3][i] = hyx; o[4][i] = hzx; o[5][i] = hzy;
            tx[i] = hxy + hxz;
            ty[i] = hyz + hyx;
            tz[i] = hzx + hzy;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) stw<W>(hs[k], b, o[k]);
        stw<W>(hx, b, tx);
        stw<W>(hy, b, ty);
        stw<W>(hz, b, tz);
#pragma unroll
        for (int i = 0; i < W; ++i) {
            ey_xm[i] = cur.ey_c[i];
            ez_xm[i] = cur.ez_c[i];
        }
        if (PF)
            cur = nxt;
        else if (more)
            load(cur, b + sx);
    }
}

template <typename T, int W, bool PF>
__global__ void __launch_bounds__(512) k_update_e_w(const InstDev<T>* __restrict__ insts,
                                                     const MemberDev* __restrict__ mem, int nmem, int total,
                                                     const int* base, int local, int io) {
    __shared__ int s_member;
    Tile t;
    const InstDev<T>* I = locate(insts, mem, nmem, total, &s_member, t, W);
    if (!I) return;
    const int ng = io ? I->negroups : 0;
    const ProbeFold<T> P{I->ebits, (const ProbeGroup<T>*)I->egroups, ng, ng ? step_of(base, local) : 0};
    if (I->coef_kind == 1)
        e_column_w<T, CoefTable<T>, W, PF>(I->F, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, I->Nx, I->Ny,
                                           I->Nz, t, P);
    else
        e_column_w<T, CoefRaw<T>, W, PF>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
}

template <typename T, int W, bool PF>
__global__ void __launch_bounds__(512) k_update_h_w(const InstDev<T>* __restrict__ insts,
                                                     const MemberDev* __restrict__ mem, int nmem, int total,
                                                     const int* base, int local, int io) {
    __shared__ int s_member;
    Tile t;
    const InstDev<T>* I = locate(insts, mem, nmem, total, &s_member, t, W);
    if (!I) return;
    const int ng = io ? I->nhgroups : 0;
    const ProbeFold<T> P{I->hbits, (const ProbeGroup<T>*)I->hgroups, ng, ng ? step_of(base, local) : 0};
    if (I->coef_kind == 1)
        h_column_w<T, CoefTable<T>, W, PF>(I->F, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, I->Nx, I->Ny,
                                           I->Nz, t, P);
    else
        h_column_w<T, CoefRaw<T>, W, PF>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
}

// ------------------------------------------------------------------------------------------------
// fused step: E and H half steps of one time step in one pass, from one field buffer into the other
//
// A workgroup owns a tile of TY rows (y) x kFz cells (z) and a range of x planes; it has TY + 2
// wavefronts of 64 lanes.  Lane l of wave w sits at y = y0 - 1 + w, z = z0 - 1 + l:
//   interior   1 <= w <= TY, 1 <= l <= kFz: E and H updated and stored;
//   lower halo w = 0 or l = 0: E computed (not stored) for the interior's H update (y-1, z-1);
//   upper halo w = TY + 1 or l = kFz + 1: H totals read for the interior's E update (y+1, z+1).
// The tile marches in x: the x+1 H splits are prefetched and carried, the x-1 E totals carried, and
// the y/z neighbours pass through LDS.  A chunk that does not start at x = 0 first computes E on
// plane x0 - 1 without storing it.  E sources are applied where they are computed (the bitmask
// marks probe cells; halo copies apply them too, only the owner records), reading the source
// samples from the pristine series; H sources and records are applied by the owner of the cell
// before its H store; monitors follow in their own launch on the written buffer.  Every value is
// formed by the same operations in the same order as in the separate E and H kernels, so results
// are identical.
// ------------------------------------------------------------------------------------------------

constexpr int kFz = 62;  // interior z cells per tile (64 lanes minus the two z halos)

template <typename T>
__device__ __forceinline__ void inject_e(const InstDev<T>& I, int64_t cell, int n, bool record, T& exy, T& exz,
                                         T& eyz, T& eyx, T& ezx, T& ezy) {
    inject_cell((const ProbeGroup<T>*)I.egroups, I.negroups, cell, n, record, exy, exz, eyz, eyx, ezx, ezy);
}

template <typename T, int TY, class CA>
__device__ __forceinline__ void fused_tile(const InstDev<T>& I, const CA& C, int x0, int x1, int y0, int z0, int n,
                                           bool io, T* sm) {
    using G = GLOBAL_T(T);
    constexpr int NT = (TY + 2) * 64;
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const int Nx = I.Nx, Ny = I.Ny, Nz = I.Nz;
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int y = y0 - 1 + w, z = z0 - 1 + l;
    const bool valid = y >= 0 && y < Ny && z >= 0 && z < Nz;
    const bool ecomp = valid && w <= TY && l <= kFz;             // E computed here
    const bool inner = ecomp && w >= 1 && l >= 1;                // E and H stored here
    const bool yin = y < Ny - 1, zin = z < Nz - 1;
    T* sHx = sm; T* sHy = sm + NT; T* sHz = sm + 2 * NT;
    T* sEx = sm + 3 * NT; T* sEy = sm + 4 * NT; T* sEz = sm + 5 * NT;
    const Fields<T>& A = I.F;
    const Fields<T>& B = I.G;
    const int xs = x0 > 0 ? x0 - 1 : x0;
    const uint32_t p = valid ? ((uint32_t)y * (uint32_t)Nz + (uint32_t)z) : 0u;
    uint32_t b = ((uint32_t)xs * (uint32_t)Ny * (uint32_t)Nz + p) * S;
    // H splits of the current plane (carried), E totals of the previous plane (carried)
    T hxy = T(0), hxz = T(0), hyz = T(0), hyx = T(0), hzx = T(0), hzy = T(0);
    if (valid) {
        hxy = ldg((const G*)A.hx_y, b); hxz = ldg((const G*)A.hx_z, b); hyz = ldg((const G*)A.hy_z, b);
        hyx = ldg((const G*)A.hy_x, b); hzx = ldg((const G*)A.hz_x, b); hzy = ldg((const G*)A.hz_y, b);
    }
    T ey_xm = T(0), ez_xm = T(0);
    // this plane's E splits and coefficient keys, loaded one plane ahead (software pipelining)
    T exy = T(0), exz = T(0), eyz = T(0), eyx = T(0), ezx = T(0), ezy = T(0);
    uint32_t ek = 0, hk = 0;
    if (ecomp) {
        exy = ldg((const G*)A.ex_y, b); exz = ldg((const G*)A.ex_z, b); eyz = ldg((const G*)A.ey_z, b);
        eyx = ldg((const G*)A.ey_x, b); ezx = ldg((const G*)A.ez_x, b); ezy = ldg((const G*)A.ez_y, b);
        ek = C.ekey(b);
        if (inner) hk = C.hkey(b);
    }
    for (int x = xs; x < x1; ++x, b += sx) {
        const bool pre = x < x0, xin = x < Nx - 1, more = x + 1 < x1;
        // next plane's H splits (E needs its Hy, Hz; the H update of the next plane needs all six)
        T nxy = T(0), nxz = T(0), nyz = T(0), nyx = T(0), nzx = T(0), nzy = T(0);
        if (valid && xin) {
            nxy = ldg((const G*)A.hx_y, b + sx); nxz = ldg((const G*)A.hx_z, b + sx);
            nyz = ldg((const G*)A.hy_z, b + sx); nyx = ldg((const G*)A.hy_x, b + sx);
            nzx = ldg((const G*)A.hz_x, b + sx); nzy = ldg((const G*)A.hz_y, b + sx);
        }
        // this plane's coefficient entries, issued before the barriers so that they arrive while the
        // tile exchanges neighbours
        ECoef<T> c{};
        HCoef<T> ch{};
        if (ecomp) c = C.e_at(ek);
        if (inner && !pre) ch = C.h_at(hk);
        // and the next plane's E splits and keys
        T pxy = T(0), pxz = T(0), pyz = T(0), pyx = T(0), pzx = T(0), pzy = T(0);
        uint32_t pek = 0, phk = 0;
        if (ecomp && more) {
            pxy = ldg((const G*)A.ex_y, b + sx); pxz = ldg((const G*)A.ex_z, b + sx);
            pyz = ldg((const G*)A.ey_z, b + sx); pyx = ldg((const G*)A.ey_x, b + sx);
            pzx = ldg((const G*)A.ez_x, b + sx); pzy = ldg((const G*)A.ez_y, b + sx);
            pek = C.ekey(b + sx);
            if (inner) phk = C.hkey(b + sx);
        }
        const T hx_c = hxy + hxz, hy_c = hyz + hyx, hz_c = hzx + hzy;
        sHx[tid] = hx_c;
        sHy[tid] = hy_c;
        sHz[tid] = hz_c;
        __syncthreads();
        T ex = T(0), ey = T(0), ez = T(0);
        if (ecomp) {
            const T hz_yn = yin ? sHz[tid + 64] : T(0), hx_yn = yin ? sHx[tid + 64] : T(0);
            const T hy_zn = zin ? sHy[tid + 1] : T(0), hx_zn = zin ? sHx[tid + 1] : T(0);
            const T hy_xn = xin ? nyz + nyx : T(0), hz_xn = xin ? nzx + nzy : T(0);
            if (yin) exy = c.ca_xy * exy + c.cb_xy * (hz_yn - hz_c);
            if (zin) exz = c.ca_xz * exz + c.cb_xz * (hy_zn - hy_c);
            if (zin) eyz = c.ca_yz * eyz + c.cb_yz * (hx_zn - hx_c);
            if (xin) eyx = c.ca_yx * eyx + c.cb_yx * (hz_xn - hz_c);
            if (xin) ezx = c.ca_zx * ezx + c.cb_zx * (hy_xn - hy_c);
            if (yin) ezy = c.ca_zy * ezy + c.cb_zy * (hx_yn - hx_c);
            if (io && I.negroups) {
                const int64_t cell = ((int64_t)x * Ny + y) * Nz + z;
                if ((I.ebits[cell >> 5] >> (cell & 31)) & 1u)
                    inject_e(I, cell, n, inner && !pre, exy, exz, eyz, eyx, ezx, ezy);
            }
            ex = exy + exz;
            ey = eyz + eyx;
            ez = ezx + ezy;
            if (inner && !pre) {
                stg((G*)B.ex_y, b, exy); stg((G*)B.ex_z, b, exz); stg((G*)B.ey_z, b, eyz);
                stg((G*)B.ey_x, b, eyx); stg((G*)B.ez_x, b, ezx); stg((G*)B.ez_y, b, ezy);
            }
        }
        sEx[tid] = ex;
        sEy[tid] = ey;
        sEz[tid] = ez;
        __syncthreads();
        if (inner && !pre) {
            const HCoef<T>& c = ch;
            const T ez_ym = y > 0 ? sEz[tid - 64] : T(0), ex_ym = y > 0 ? sEx[tid - 64] : T(0);
            const T ey_zm = z > 0 ? sEy[tid - 1] : T(0), ex_zm = z > 0 ? sEx[tid - 1] : T(0);
            T xy = c.da_xy * hxy + (c.db_xy2 * ez - c.db_xy1 * ez_ym);
            T xz = c.da_xz * hxz + (c.db_xz2 * ey - c.db_xz1 * ey_zm);
            T yz = c.da_yz * hyz + (c.db_yz2 * ex - c.db_yz1 * ex_zm);
            T yx = c.da_yx * hyx + (c.db_yx2 * ez - c.db_yx1 * ez_xm);
            T zx = c.da_zx * hzx + (c.db_zx2 * ey - c.db_zx1 * ey_xm);
            T zy = c.da_zy * hzy + (c.db_zy2 * ex - c.db_zy1 * ex_ym);
            if (io && I.nhgroups) {  // H sources and records where H is computed
                const int64_t cell = ((int64_t)x * Ny + y) * Nz + z;
                if ((I.hbits[cell >> 5] >> (cell & 31)) & 1u)
                    inject_cell((const ProbeGroup<T>*)I.hgroups, I.nhgroups, cell, n, true, xy, xz, yz, yx, zx, zy);
            }
            stg((G*)B.hx_y, b, xy); stg((G*)B.hx_z, b, xz); stg((G*)B.hy_z, b, yz);
            stg((G*)B.hy_x, b, yx); stg((G*)B.hz_x, b, zx); stg((G*)B.hz_y, b, zy);
        }
        ey_xm = ey;
        ez_xm = ez;
        hxy = nxy; hxz = nxz; hyz = nyz; hyx = nyx; hzx = nzx; hzy = nzy;
        exy = pxy; exz = pxz; eyz = pyz; eyx = pyx; ezx = pzx; ezy = pzy;
        ek = pek;
        hk = phk;
    }
}

// The same pass with E one plane ahead of H, so that one barrier per plane suffices: iteration i
// writes plane i's H totals to LDS, waits once, then updates H on plane i - 1 (from plane i - 1's E
// totals, written to LDS in the previous iteration) and E on plane i (from the H totals just
// exchanged).  The two LDS exchanges alternate between two buffers by plane parity, so a wave
// writing plane i + 1's totals never overwrites what a slower wave still reads.  H splits are
// loaded two planes ahead, E splits and coefficient values one plane ahead and coefficient keys two
// planes ahead, so that a table entry's gather is in flight for a whole plane.  Every value is
// formed as in fused_tile.
template <typename T>
struct Split6 {
    T xy, xz, yz, yx, zx, zy;
};

template <typename T, int TY, class CA>
__device__ __forceinline__ void fused_tile_skew(const InstDev<T>& I, const CA& C, int x0, int x1, int y0, int z0,
                                                int n, bool io, T* sm) {
    using G = GLOBAL_T(T);
    constexpr int NT = (TY + 2) * 64;
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const int Nx = I.Nx, Ny = I.Ny, Nz = I.Nz;
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int y = y0 - 1 + w, z = z0 - 1 + l;
    const bool valid = y >= 0 && y < Ny && z >= 0 && z < Nz;
    const bool ecomp = valid && w <= TY && l <= kFz;  // E computed here
    const bool inner = ecomp && w >= 1 && l >= 1;     // E and H stored here
    const bool yin = y < Ny - 1, zin = z < Nz - 1;
    T* const sH = sm;           // [parity][component][NT] H totals
    T* const sE = sm + 6 * NT;  // [parity][component][NT] E totals
    const Fields<T>& A = I.F;
    const Fields<T>& B = I.G;
    const int xs = x0 > 0 ? x0 - 1 : x0;  // a chunk not starting at 0 computes E on x0 - 1 first
    const int xh = min(x1, Nx - 1);       // last plane whose H splits are read
    const uint32_t p = valid ? ((uint32_t)y * (uint32_t)Nz + (uint32_t)z) : 0u;
    const uint32_t b0 = ((uint32_t)xs * (uint32_t)Ny * (uint32_t)Nz + p) * S;
    auto off = [&](int x) { return b0 + (uint32_t)(x - xs) * sx; };
    auto load_h = [&](int x) {
        Split6<T> h{};
        if (valid && x <= xh) {
            const uint32_t b = off(x);
            h.xy = ldg((const G*)A.hx_y, b); h.xz = ldg((const G*)A.hx_z, b); h.yz = ldg((const G*)A.hy_z, b);
            h.yx = ldg((const G*)A.hy_x, b); h.zx = ldg((const G*)A.hz_x, b); h.zy = ldg((const G*)A.hz_y, b);
        }
        return h;
    };
    auto load_e = [&](int x) {
        Split6<T> e{};
        const uint32_t b = off(x);
        e.xy = ldg((const G*)A.ex_y, b); e.xz = ldg((const G*)A.ex_z, b); e.yz = ldg((const G*)A.ey_z, b);
        e.yx = ldg((const G*)A.ey_x, b); e.zx = ldg((const G*)A.ez_x, b); e.zy = ldg((const G*)A.ez_y, b);
        return e;
    };
    Split6<T> hA{}, hB = load_h(xs), hC = load_h(xs + 1);  // H splits of planes i - 1, i, i + 1
    Split6<T> e{};                                          // E splits of plane i
    ECoef<T> c{};                                           // E coefficients of plane i
    HCoef<T> ch{};                                          // H coefficients of plane i - 1
    uint32_t ek = 0, hk = 0;                                // keys: E of plane i + 1, H of plane i
    if (ecomp) {
        e = load_e(xs);
        c = C.e_at(C.ekey(off(xs)));
        if (xs + 1 < x1) ek = C.ekey(off(xs + 1));
    }
    if (inner && xs >= x0 && xs < x1) hk = C.hkey(off(xs));
    T ex1 = T(0), ey1 = T(0), ez1 = T(0);  // E totals of plane i - 1 at this cell
    T ey2 = T(0), ez2 = T(0);              // E totals of plane i - 2
    for (int i = xs; i <= x1; ++i) {
        const bool eplane = i < x1;     // E updated on plane i
        const bool hplane = i - 1 >= x0;  // H updated on plane i - 1 (not the extra plane x0 - 1)
        const int par = i & 1;
        // the next iteration's loads: coefficient values (E of plane i + 1, H of plane i) from the
        // keys loaded an iteration earlier, then splits and the keys after them
        ECoef<T> cN{};
        HCoef<T> chN{};
        if (ecomp && i + 1 < x1) cN = C.e_at(ek);
        if (inner && i >= x0 && i < x1) chN = C.h_at(hk);
        const Split6<T> hN = load_h(i + 2);
        Split6<T> eN{};
        uint32_t ekN = 0, hkN = 0;
        if (ecomp && i + 1 < x1) eN = load_e(i + 1);
        if (ecomp && i + 2 < x1) ekN = C.ekey(off(i + 2));
        if (inner && i + 1 >= x0 && i + 1 < x1) hkN = C.hkey(off(i + 1));
        const T hx_c = hB.xy + hB.xz, hy_c = hB.yz + hB.yx, hz_c = hB.zx + hB.zy;
        if (eplane) {
            sH[(par * 3 + 0) * NT + tid] = hx_c;
            sH[(par * 3 + 1) * NT + tid] = hy_c;
            sH[(par * 3 + 2) * NT + tid] = hz_c;
        }
        __syncthreads();
        if (hplane && inner) {
            const int pp = par ^ 1;
            const T* sEx = sE + (pp * 3 + 0) * NT;
            const T* sEy = sE + (pp * 3 + 1) * NT;
            const T* sEz = sE + (pp * 3 + 2) * NT;
            const T ex = ex1, ey = ey1, ez = ez1, ey_xm = ey2, ez_xm = ez2;
            const T ez_ym = y > 0 ? sEz[tid - 64] : T(0), ex_ym = y > 0 ? sEx[tid - 64] : T(0);
            const T ey_zm = z > 0 ? sEy[tid - 1] : T(0), ex_zm = z > 0 ? sEx[tid - 1] : T(0);
            T xy = ch.da_xy * hA.xy + (ch.db_xy2 * ez - ch.db_xy1 * ez_ym);
            T xz = ch.da_xz * hA.xz + (ch.db_xz2 * ey - ch.db_xz1 * ey_zm);
            T yz = ch.da_yz * hA.yz + (ch.db_yz2 * ex - ch.db_yz1 * ex_zm);
            T yx = ch.da_yx * hA.yx + (ch.db_yx2 * ez - ch.db_yx1 * ez_xm);
            T zx = ch.da_zx * hA.zx + (ch.db_zx2 * ey - ch.db_zx1 * ey_xm);
            T zy = ch.da_zy * hA.zy + (ch.db_zy2 * ex - ch.db_zy1 * ex_ym);
            if (io && I.nhgroups) {  // H sources and records where H is computed
                const int64_t cell = ((int64_t)(i - 1) * Ny + y) * Nz + z;
                if ((I.hbits[cell >> 5] >> (cell & 31)) & 1u)
                    inject_cell((const ProbeGroup<T>*)I.hgroups, I.nhgroups, cell, n, true, xy, xz, yz, yx, zx, zy);
            }
            const uint32_t b = off(i - 1);
            stg((G*)B.hx_y, b, xy); stg((G*)B.hx_z, b, xz); stg((G*)B.hy_z, b, yz);
            stg((G*)B.hy_x, b, yx); stg((G*)B.hz_x, b, zx); stg((G*)B.hz_y, b, zy);
        }
        if (eplane) {
            T ex = T(0), ey = T(0), ez = T(0);
            if (ecomp) {
                const T* sHx = sH + (par * 3 + 0) * NT;
                const T* sHy = sH + (par * 3 + 1) * NT;
                const T* sHz = sH + (par * 3 + 2) * NT;
                const bool xin = i < Nx - 1;
                const T hz_yn = yin ? sHz[tid + 64] : T(0), hx_yn = yin ? sHx[tid + 64] : T(0);
                const T hy_zn = zin ? sHy[tid + 1] : T(0), hx_zn = zin ? sHx[tid + 1] : T(0);
                const T hy_xn = xin ? hC.yz + hC.yx : T(0), hz_xn = xin ? hC.zx + hC.zy : T(0);
                T exy = e.xy, exz = e.xz, eyz = e.yz, eyx = e.yx, ezx = e.zx, ezy = e.zy;
                if (yin) exy = c.ca_xy * exy + c.cb_xy * (hz_yn - hz_c);
                if (zin) exz = c.ca_xz * exz + c.cb_xz * (hy_zn - hy_c);
                if (zin) eyz = c.ca_yz * eyz + c.cb_yz * (hx_zn - hx_c);
                if (xin) eyx = c.ca_yx * eyx + c.cb_yx * (hz_xn - hz_c);
                if (xin) ezx = c.ca_zx * ezx + c.cb_zx * (hy_xn - hy_c);
                if (yin) ezy = c.ca_zy * ezy + c.cb_zy * (hx_yn - hx_c);
                const bool own = inner && i >= x0;
                if (io && I.negroups) {
                    const int64_t cell = ((int64_t)i * Ny + y) * Nz + z;
                    if ((I.ebits[cell >> 5] >> (cell & 31)) & 1u)
                        inject_e(I, cell, n, own, exy, exz, eyz, eyx, ezx, ezy);
                }
                ex = exy + exz;
                ey = eyz + eyx;
                ez = ezx + ezy;
                if (own) {
                    const uint32_t b = off(i);
                    stg((G*)B.ex_y, b, exy); stg((G*)B.ex_z, b, exz); stg((G*)B.ey_z, b, eyz);
                    stg((G*)B.ey_x, b, eyx); stg((G*)B.ez_x, b, ezx); stg((G*)B.ez_y, b, ezy);
                }
            }
            sE[(par * 3 + 0) * NT + tid] = ex;
            sE[(par * 3 + 1) * NT + tid] = ey;
            sE[(par * 3 + 2) * NT + tid] = ez;
            ey2 = ey1;
            ez2 = ez1;
            ex1 = ex;
            ey1 = ey;
            ez1 = ez;
        }
        hA = hB;
        hB = hC;
        hC = hN;
        e = eN;
        c = cN;
        ch = chN;
        ek = ekN;
        hk = hkN;
    }
}

template <typename T, int TY, bool D>
__global__ void __launch_bounds__((TY + 2) * 64) k_step_fused(const InstDev<T>* __restrict__ insts,
                                                              const MemberDev* __restrict__ mem, int nmem, int total,
                                                              const int* base, int local, int io) {
    __shared__ int s_member;
    __shared__ T sm[(D ? 12 : 6) * (TY + 2) * 64];
    const int L = xcd_logical(blockIdx.x, gridDim.x);
    if (threadIdx.x == 0) {
        int m = 0;
        if (L < total)
            while (m + 1 < nmem && mem[m + 1].tile0 <= L) ++m;
        s_member = m;
    }
    __syncthreads();
    if (L >= total) return;  // uniform over the workgroup: no barrier is skipped by part of it
    const int m = __builtin_amdgcn_readfirstlane(s_member);
    const MemberDev M = mem[m];
    const int local_t = L - M.tile0;
    const int e = local_t % M.K, tile = local_t / M.K;
    const InstDev<T>* I = insts + __builtin_amdgcn_readfirstlane(M.inst0 + e);
    const int tyz = I->npb, tz_n = I->tz_n;
    const int pb = tile % tyz, xc = tile / tyz;
    const int y0 = (pb / tz_n) * TY, z0 = (pb % tz_n) * kFz;
    const int x0 = xc * I->xplanes, x1 = min(I->Nx, x0 + I->xplanes);
    const int n = step_of(base, local);
    if (D) {  // E one plane ahead of H, one barrier per plane
        if (I->coef_kind == 1)
            fused_tile_skew<T, TY, CoefTable<T>>(*I, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, x0, x1, y0, z0,
                                                 n, io != 0, sm);
        else
            fused_tile_skew<T, TY, CoefRaw<T>>(*I, CoefRaw<T>{I->raw}, x0, x1, y0, z0, n, io != 0, sm);
    } else {
        if (I->coef_kind == 1)
            fused_tile<T, TY, CoefTable<T>>(*I, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, x0, x1, y0, z0, n,
                                            io != 0, sm);
        else
            fused_tile<T, TY, CoefRaw<T>>(*I, CoefRaw<T>{I->raw}, x0, x1, y0, z0, n, io != 0, sm);
    }
}

// ------------------------------------------------------------------------------------------------
// adjoint (reverse-time) kernels
//
// Forward step n (state e = 6 E splits, h = 6 H splits; totals are sums of splits):
//   e' = Ca e + Cb D(H^n)          e'' = e' + sources          y_E = E''(probe cells)
//   h' = Da h + G(E'')             h'' = h' + sources          y_H = H''(probe cells)
// For an objective J(y) with g = dJ/dy, the reverse step n, given the split adjoints of the state
// after the step (lambda_h in the H arrays; in the E arrays nu of step n+1, whose Ca scaling is
// applied on reading), is:
//   mu  = lambda_h + g_H (both splits of each probed H component)           k_adj_inject (H)
//   nu  = Ca nu_old + S^T G^T mu                                             k_adj_e
//   nu += g_E (both splits)                                                  k_adj_inject (E)
//   dJ/dCb += nu * D(H^n)   (design box, forward H^n saved or recomputed)    k_adj_h's extra workgroups
//   lambda_h = Da mu + S^T D^T (Cb nu)                                       k_adj_h
// where S^T copies a total's adjoint into both of its splits.  Split updates that the forward step
// skips (upper edges of the GPU packing) pass their adjoint through unchanged.
// ------------------------------------------------------------------------------------------------

enum {  // coefficient positions
    C_CA_XY = 0, C_CA_XZ, C_CB_XY, C_CB_XZ, C_CA_YZ, C_CA_YX, C_CB_YZ, C_CB_YX, C_CA_ZX, C_CA_ZY, C_CB_ZX, C_CB_ZY,
    C_DA_XY, C_DA_XZ, C_DB_XY1, C_DB_XY2, C_DB_XZ1, C_DB_XZ2, C_DA_YZ, C_DA_YX, C_DB_YZ1, C_DB_YZ2, C_DB_YX1,
    C_DB_YX2, C_DA_ZX, C_DA_ZY, C_DB_ZX1, C_DB_ZX2, C_DB_ZY1, C_DB_ZY2
};

// The total of a component at one cell, formed from its splits as the update kernels form it.
template <typename T>
__device__ __forceinline__ T total_at(const InstDev<T>& I, int field, int64_t i) {
    switch (field) {
        case EX: return I.F.ex_y[i] + I.F.ex_z[i];
        case EY: return I.F.ey_z[i] + I.F.ey_x[i];
        case EZ: return I.F.ez_x[i] + I.F.ez_y[i];
        case HX: return I.F.hx_y[i] + I.F.hx_z[i];
        case HY: return I.F.hy_z[i] + I.F.hy_x[i];
        default: return I.F.hz_x[i] + I.F.hz_y[i];
    }
}

// Batched flat launches: item i (an instance, a design box) owns blocks [blk0[i], blk0[i+1]).
__device__ __forceinline__ int item_of(const int64_t* __restrict__ blk0, int n, int64_t b) {
    int i = 0;
    while (i + 1 < n && blk0[i + 1] <= b) ++i;
    return i;
}

// Reverse E half step over one column of x planes (the transpose of h_column followed by the
// Ca scaling of the next step's e_column).  Marches x upward and carries the x+1 plane's H
// coefficients and the two mu splits that read it.
template <typename T, class CA>
__device__ __forceinline__ void adj_e_column(const Fields<T>& F, const CA& C, int Nx, int Ny, int Nz, const Tile& t,
                                             const AdjFold<T>& P) {
    using G = GLOBAL_T(T);
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S, sy = (uint32_t)Nz * S;
    const G* mxy = (const G*)F.hx_y; const G* mxz = (const G*)F.hx_z;
    const G* myz = (const G*)F.hy_z; const G* myx = (const G*)F.hy_x;
    const G* mzx = (const G*)F.hz_x; const G* mzy = (const G*)F.hz_y;
    G* nxy = (G*)F.ex_y; G* nxz = (G*)F.ex_z; G* nyz = (G*)F.ey_z;
    G* nyx = (G*)F.ey_x; G* nzx = (G*)F.ez_x; G* nzy = (G*)F.ez_y;
    const bool yin = t.y < Ny - 1, zin = t.z < Nz - 1;
    if (t.x0 >= t.x1) return;
    uint32_t b = ((uint32_t)t.x0 * (uint32_t)Ny * (uint32_t)Nz + (uint32_t)t.p) * S;
    HCoef<T> hc = C.h(b);
    T m_yx = ldg(myx, b), m_zx = ldg(mzx, b);
    for (int x = t.x0; x < t.x1; ++x, b += sx) {
        const bool xin = x < Nx - 1;
        HCoef<T> hn{};
        T n_yx = T(0), n_zx = T(0);
        if (xin) {
            hn = C.h(b + sx);
            n_yx = ldg(myx, b + sx);
            n_zx = ldg(mzx, b + sx);
        }
        const T m_xy = ldg(mxy, b), m_xz = ldg(mxz, b), m_yz = ldg(myz, b), m_zy = ldg(mzy, b);
        // E totals' adjoints: every H split that reads E at this cell (as e or as e- of a neighbour)
        T ez = hc.db_xy2 * m_xy + hc.db_yx2 * m_yx;
        if (yin) ez = ez - C.at(b + sy, C_DB_XY1) * ldg(mxy, b + sy);
        if (xin) ez = ez - hn.db_yx1 * n_yx;
        T ey = hc.db_xz2 * m_xz + hc.db_zx2 * m_zx;
        if (zin) ey = ey - C.at(b + S, C_DB_XZ1) * ldg(mxz, b + S);
        if (xin) ey = ey - hn.db_zx1 * n_zx;
        T ex = hc.db_yz2 * m_yz + hc.db_zy2 * m_zy;
        if (zin) ex = ex - C.at(b + S, C_DB_YZ1) * ldg(myz, b + S);
        if (yin) ex = ex - C.at(b + sy, C_DB_ZY1) * ldg(mzy, b + sy);
        const ECoef<T> ec = C.e(b);
        T o_xy = (yin ? ec.ca_xy : T(1)) * ldg(nxy, b) + ex;
        T o_xz = (zin ? ec.ca_xz : T(1)) * ldg(nxz, b) + ex;
        T o_yz = (zin ? ec.ca_yz : T(1)) * ldg(nyz, b) + ey;
        T o_yx = (xin ? ec.ca_yx : T(1)) * ldg(nyx, b) + ey;
        T o_zx = (xin ? ec.ca_zx : T(1)) * ldg(nzx, b) + ez;
        T o_zy = (yin ? ec.ca_zy : T(1)) * ldg(nzy, b) + ez;
        P.apply(b / S, o_xy, o_xz, o_yz, o_yx, o_zx, o_zy);
        stg(nxy, b, o_xy); stg(nxz, b, o_xz); stg(nyz, b, o_yz);
        stg(nyx, b, o_yx); stg(nzx, b, o_zx); stg(nzy, b, o_zy);
        hc = hn;
        m_yx = n_yx;
        m_zx = n_zx;
    }
}

// Reverse H half step over one column (the transpose of e_column followed by the Da scaling of the
// next step's h_column).  Marches x upward and carries Cb * nu of the x-1 plane's two x splits.
template <typename T, class CA>
__device__ __forceinline__ void adj_h_column(const Fields<T>& F, const CA& C, int Nx, int Ny, int Nz, const Tile& t,
                                             const AdjFold<T>& P) {
    using G = GLOBAL_T(T);
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S, sy = (uint32_t)Nz * S;
    const G* nxy = (const G*)F.ex_y; const G* nxz = (const G*)F.ex_z; const G* nyz = (const G*)F.ey_z;
    const G* nyx = (const G*)F.ey_x; const G* nzx = (const G*)F.ez_x; const G* nzy = (const G*)F.ez_y;
    G* mxy = (G*)F.hx_y; G* mxz = (G*)F.hx_z; G* myz = (G*)F.hy_z;
    G* myx = (G*)F.hy_x; G* mzx = (G*)F.hz_x; G* mzy = (G*)F.hz_y;
    const bool yin = t.y < Ny - 1, zin = t.z < Nz - 1, ylo = t.y > 0, zlo = t.z > 0;
    if (t.x0 >= t.x1) return;
    uint32_t b = ((uint32_t)t.x0 * (uint32_t)Ny * (uint32_t)Nz + (uint32_t)t.p) * S;
    T p_yx = T(0), p_zx = T(0);
    if (t.x0 > 0) {
        p_yx = C.at(b - sx, C_CB_YX) * ldg(nyx, b - sx);
        p_zx = C.at(b - sx, C_CB_ZX) * ldg(nzx, b - sx);
    }
    for (int x = t.x0; x < t.x1; ++x, b += sx) {
        const bool xin = x < Nx - 1, xlo = x > 0;
        const ECoef<T> ec = C.e(b);
        const T v_xy = ldg(nxy, b), v_xz = ldg(nxz, b), v_yz = ldg(nyz, b);
        const T v_yx = ldg(nyx, b), v_zx = ldg(nzx, b), v_zy = ldg(nzy, b);
        // H totals' adjoints: every E split whose curl difference reads H at this cell
        T hz = T(0), hy = T(0), hx = T(0);
        if (yin) hz = hz - ec.cb_xy * v_xy;
        if (ylo) hz = hz + C.at(b - sy, C_CB_XY) * ldg(nxy, b - sy);
        if (xin) hz = hz - ec.cb_yx * v_yx;
        if (xlo) hz = hz + p_yx;
        if (zin) hy = hy - ec.cb_xz * v_xz;
        if (zlo) hy = hy + C.at(b - S, C_CB_XZ) * ldg(nxz, b - S);
        if (xin) hy = hy - ec.cb_zx * v_zx;
        if (xlo) hy = hy + p_zx;
        if (zin) hx = hx - ec.cb_yz * v_yz;
        if (zlo) hx = hx + C.at(b - S, C_CB_YZ) * ldg(nyz, b - S);
        if (yin) hx = hx - ec.cb_zy * v_zy;
        if (ylo) hx = hx + C.at(b - sy, C_CB_ZY) * ldg(nzy, b - sy);
        const HCoef<T> hc = C.h(b);
        T o_xy = hc.da_xy * ldg(mxy, b) + hx;
        T o_xz = hc.da_xz * ldg(mxz, b) + hx;
        T o_yz = hc.da_yz * ldg(myz, b) + hy;
        T o_yx = hc.da_yx * ldg(myx, b) + hy;
        T o_zx = hc.da_zx * ldg(mzx, b) + hz;
        T o_zy = hc.da_zy * ldg(mzy, b) + hz;
        P.apply(b / S, o_xy, o_xz, o_yz, o_yx, o_zx, o_zy);
        stg(mxy, b, o_xy); stg(mxz, b, o_xz); stg(myz, b, o_yz);
        stg(myx, b, o_yx); stg(mzx, b, o_zx); stg(mzy, b, o_zy);
        p_yx = ec.cb_yx * v_yx;
        p_zx = ec.cb_zx * v_zx;
    }
}

// The same two reverse half steps software-pipelined: the coefficient entry keys of plane x + 2 and
// every other load of plane x + 1 are issued before plane x's arithmetic and stores, so a lane keeps
// two planes of loads in flight and never waits on a key just loaded.  Arithmetic (and so every
// result bit) is that of adj_e_column / adj_h_column.
template <typename T>
struct AEIn {
    T m_xy, m_xz, m_yz, m_zy;              // mu at the cell
    T n_yx, n_zx;                          // mu x splits at x + 1
    T m_xy_y, m_zy_y, m_xz_z, m_yz_z;      // mu at y + 1 and z + 1
    T d_xy1_y, d_zy1_y, d_xz1_z, d_yz1_z;  // their Db coefficients
    HCoef<T> hn;                           // H coefficients at x + 1
    T ca_xy, ca_xz, ca_yz, ca_yx, ca_zx, ca_zy;
    T v_xy, v_xz, v_yz, v_yx, v_zx, v_zy;  // nu at the cell
    uint32_t pw;                           // adjoint-source bitmask word of the cell
};
struct AEKeys {
    uint32_t hx, hy, hz, e;  // H entry at x + 1, y + 1, z + 1; E entry at the cell
};

template <typename T, class CA>
__device__ __forceinline__ void adj_e_column_pf(const Fields<T>& F, const CA& C, int Nx, int Ny, int Nz, const Tile& t,
                                                const AdjFold<T>& P) {
    using G = GLOBAL_T(T);
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S, sy = (uint32_t)Nz * S;
    const G* mxy = (const G*)F.hx_y; const G* mxz = (const G*)F.hx_z;
    const G* myz = (const G*)F.hy_z; const G* myx = (const G*)F.hy_x;
    const G* mzx = (const G*)F.hz_x; const G* mzy = (const G*)F.hz_y;
    G* nxy = (G*)F.ex_y; G* nxz = (G*)F.ex_z; G* nyz = (G*)F.ey_z;
    G* nyx = (G*)F.ey_x; G* nzx = (G*)F.ez_x; G* nzy = (G*)F.ez_y;
    const bool yin = t.y < Ny - 1, zin = t.z < Nz - 1;
    if (t.x0 >= t.x1) return;
    auto keys = [&](uint32_t b, bool xin) {
        AEKeys k;
        k.hx = xin ? C.hkey(b + sx) : 0u;
        k.hy = yin ? C.hkey(b + sy) : 0u;
        k.hz = zin ? C.hkey(b + S) : 0u;
        k.e = C.ekey(b);
        return k;
    };
    auto load = [&](AEIn<T>& v, uint32_t b, bool xin, const AEKeys& k) {
        v.m_xy = ldg(mxy, b); v.m_xz = ldg(mxz, b); v.m_yz = ldg(myz, b); v.m_zy = ldg(mzy, b);
        v.n_yx = xin ? ldg(myx, b + sx) : T(0);
        v.n_zx = xin ? ldg(mzx, b + sx) : T(0);
        v.m_xy_y = yin ? ldg(mxy, b + sy) : T(0);
        v.m_zy_y = yin ? ldg(mzy, b + sy) : T(0);
        v.m_xz_z = zin ? ldg(mxz, b + S) : T(0);
        v.m_yz_z = zin ? ldg(myz, b + S) : T(0);
        v.v_xy = ldg((const G*)nxy, b); v.v_xz = ldg((const G*)nxz, b); v.v_yz = ldg((const G*)nyz, b);
        v.v_yx = ldg((const G*)nyx, b); v.v_zx = ldg((const G*)nzx, b); v.v_zy = ldg((const G*)nzy, b);
        v.d_xy1_y = yin ? C.at_key(k.hy, C_DB_XY1) : T(0);
        v.d_zy1_y = yin ? C.at_key(k.hy, C_DB_ZY1) : T(0);
        v.d_xz1_z = zin ? C.at_key(k.hz, C_DB_XZ1) : T(0);
        v.d_yz1_z = zin ? C.at_key(k.hz, C_DB_YZ1) : T(0);
        if (xin) v.hn = C.h_at(k.hx);
        else v.hn = HCoef<T>{};
        const ECoef<T> ec = C.e_at(k.e);
        v.ca_xy = ec.ca_xy; v.ca_xz = ec.ca_xz; v.ca_yz = ec.ca_yz;
        v.ca_yx = ec.ca_yx; v.ca_zx = ec.ca_zx; v.ca_zy = ec.ca_zy;
        v.pw = P.word(b / S);
    };
    uint32_t b = ((uint32_t)t.x0 * (uint32_t)Ny * (uint32_t)Nz + (uint32_t)t.p) * S;
    HCoef<T> hc = C.h(b);
    T m_yx = ldg(myx, b), m_zx = ldg(mzx, b);
    AEIn<T> cur;
    load(cur, b, t.x0 < Nx - 1, keys(b, t.x0 < Nx - 1));
    AEKeys kn{};
    if (t.x0 + 1 < t.x1) kn = keys(b + sx, t.x0 + 1 < Nx - 1);
    for (int x = t.x0; x < t.x1; ++x, b += sx) {
        const bool xin = x < Nx - 1, more = x + 1 < t.x1;
        AEIn<T> nxt;
        if (more) {
            load(nxt, b + sx, x + 1 < Nx - 1, kn);
            if (x + 2 < t.x1) kn = keys(b + 2u * sx, x + 2 < Nx - 1);
        }
        T ez = hc.db_xy2 * cur.m_xy + hc.db_yx2 * m_yx;
        if (yin) ez = ez - cur.d_xy1_y * cur.m_xy_y;
        if (xin) ez = ez - cur.hn.db_yx1 * cur.n_yx;
        T ey = hc.db_xz2 * cur.m_xz + hc.db_zx2 * m_zx;
        if (zin) ey = ey - cur.d_xz1_z * cur.m_xz_z;
        if (xin) ey = ey - cur.hn.db_zx1 * cur.n_zx;
        T ex = hc.db_yz2 * cur.m_yz + hc.db_zy2 * cur.m_zy;
        if (zin) ex = ex - cur.d_yz1_z * cur.m_yz_z;
        if (yin) ex = ex - cur.d_zy1_y * cur.m_zy_y;
        T o_xy = (yin ? cur.ca_xy : T(1)) * cur.v_xy + ex;
        T o_xz = (zin ? cur.ca_xz : T(1)) * cur.v_xz + ex;
        T o_yz = (zin ? cur.ca_yz : T(1)) * cur.v_yz + ey;
        T o_yx = (xin ? cur.ca_yx : T(1)) * cur.v_yx + ey;
        T o_zx = (xin ? cur.ca_zx : T(1)) * cur.v_zx + ez;
        T o_zy = (yin ? cur.ca_zy : T(1)) * cur.v_zy + ez;
        P.apply(b / S, cur.pw, o_xy, o_xz, o_yz, o_yx, o_zx, o_zy);
        stg(nxy, b, o_xy); stg(nxz, b, o_xz); stg(nyz, b, o_yz);
        stg(nyx, b, o_yx); stg(nzx, b, o_zx); stg(nzy, b, o_zy);
        hc = cur.hn;
        m_yx = cur.n_yx;
        m_zx = cur.n_zx;
        if (more) cur = nxt;
    }
}

template <typename T>
struct AHIn {
    T cb_xy, cb_xz, cb_yz, cb_yx, cb_zx, cb_zy;  // E coefficients at the cell
    T v_xy, v_xz, v_yz, v_yx, v_zx, v_zy;        // nu at the cell
    T w_xy_y, w_zy_y, w_xz_z, w_yz_z;            // Cb * nu at y - 1 and z - 1
    T da_xy, da_xz, da_yz, da_yx, da_zx, da_zy;  // H coefficients at the cell
    T m_xy, m_xz, m_yz, m_yx, m_zx, m_zy;        // mu at the cell
    uint32_t pw;                                 // adjoint-source bitmask word of the cell
};
struct AHKeys {
    uint32_t e, ey, ez, h;  // E entry at the cell, y - 1, z - 1; H entry at the cell
};

template <typename T, class CA>
__device__ __forceinline__ void adj_h_column_pf(const Fields<T>& F, const CA& C, int Nx, int Ny, int Nz, const Tile& t,
                                                const AdjFold<T>& P) {
    using G = GLOBAL_T(T);
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S, sy = (uint32_t)Nz * S;
    const G* nxy = (const G*)F.ex_y; const G* nxz = (const G*)F.ex_z; const G* nyz = (const G*)F.ey_z;
    const G* nyx = (const G*)F.ey_x; const G* nzx = (const G*)F.ez_x; const G* nzy = (const G*)F.ez_y;
    G* mxy = (G*)F.hx_y; G* mxz = (G*)F.hx_z; G* myz = (G*)F.hy_z;
    G* myx = (G*)F.hy_x; G* mzx = (G*)F.hz_x; G* mzy = (G*)F.hz_y;
    const bool yin = t.y < Ny - 1, zin = t.z < Nz - 1, ylo = t.y > 0, zlo = t.z > 0;
    if (t.x0 >= t.x1) return;
    auto keys = [&](uint32_t b) {
        AHKeys k;
        k.e = C.ekey(b);
        k.ey = ylo ? C.ekey(b - sy) : 0u;
        k.ez = zlo ? C.ekey(b - S) : 0u;
        k.h = C.hkey(b);
        return k;
    };
    // the neighbour products are formed as the plain kernel forms them (coefficient * nu)
    auto load = [&](AHIn<T>& v, uint32_t b, const AHKeys& k) {
        v.v_xy = ldg(nxy, b); v.v_xz = ldg(nxz, b); v.v_yz = ldg(nyz, b);
        v.v_yx = ldg(nyx, b); v.v_zx = ldg(nzx, b); v.v_zy = ldg(nzy, b);
        v.m_xy = ldg((const G*)mxy, b); v.m_xz = ldg((const G*)mxz, b); v.m_yz = ldg((const G*)myz, b);
        v.m_yx = ldg((const G*)myx, b); v.m_zx = ldg((const G*)mzx, b); v.m_zy = ldg((const G*)mzy, b);
        T a_xy = T(0), a_zy = T(0), a_xz = T(0), a_yz = T(0), c_xy = T(0), c_zy = T(0), c_xz = T(0), c_yz = T(0);
        if (ylo) {
            a_xy = ldg(nxy, b - sy); a_zy = ldg(nzy, b - sy);
            c_xy = C.at_key(k.ey, C_CB_XY); c_zy = C.at_key(k.ey, C_CB_ZY);
        }
        if (zlo) {
            a_xz = ldg(nxz, b - S); a_yz = ldg(nyz, b - S);
            c_xz = C.at_key(k.ez, C_CB_XZ); c_yz = C.at_key(k.ez, C_CB_YZ);
        }
        const ECoef<T> ec = C.e_at(k.e);
        v.cb_xy = ec.cb_xy; v.cb_xz = ec.cb_xz; v.cb_yz = ec.cb_yz;
        v.cb_yx = ec.cb_yx; v.cb_zx = ec.cb_zx; v.cb_zy = ec.cb_zy;
        const HCoef<T> hc = C.h_at(k.h);
        v.da_xy = hc.da_xy; v.da_xz = hc.da_xz; v.da_yz = hc.da_yz;
        v.da_yx = hc.da_yx; v.da_zx = hc.da_zx; v.da_zy = hc.da_zy;
        v.w_xy_y = c_xy * a_xy; v.w_zy_y = c_zy * a_zy;
        v.w_xz_z = c_xz * a_xz; v.w_yz_z = c_yz * a_yz;
        v.pw = P.word(b / S);
    };
    uint32_t b = ((uint32_t)t.x0 * (uint32_t)Ny * (uint32_t)Nz + (uint32_t)t.p) * S;
    T p_yx = T(0), p_zx = T(0);
    if (t.x0 > 0) {
        p_yx = C.at(b - sx, C_CB_YX) * ldg(nyx, b - sx);
        p_zx = C.at(b - sx, C_CB_ZX) * ldg(nzx, b - sx);
    }
    AHIn<T> cur;
    load(cur, b, keys(b));
    AHKeys kn{};
    if (t.x0 + 1 < t.x1) kn = keys(b + sx);
    for (int x = t.x0; x < t.x1; ++x, b += sx) {
        const bool xin = x < Nx - 1, xlo = x > 0, more = x + 1 < t.x1;
        AHIn<T> nxt;
        if (more) {
            load(nxt, b + sx, kn);
            if (x + 2 < t.x1) kn = keys(b + 2u * sx);
        }
        T hz = T(0), hy = T(0), hx = T(0);
        if (yin) hz = hz - cur.cb_xy * cur.v_xy;
        if (ylo) hz = hz + cur.w_xy_y;
        if (xin) hz = hz - cur.cb_yx * cur.v_yx;
        if (xlo) hz = hz + p_yx;
        if (zin) hy = hy - cur.cb_xz * cur.v_xz;
        if (zlo) hy = hy + cur.w_xz_z;
        if (xin) hy = hy - cur.cb_zx * cur.v_zx;
        if (xlo) hy = hy + p_zx;
        if (zin) hx = hx - cur.cb_yz * cur.v_yz;
        if (zlo) hx = hx + cur.w_yz_z;
        if (yin) hx = hx - cur.cb_zy * cur.v_zy;
        if (ylo) hx = hx + cur.w_zy_y;
        T o_xy = cur.da_xy * cur.m_xy + hx;
        T o_xz = cur.da_xz * cur.m_xz + hx;
        T o_yz = cur.da_yz * cur.m_yz + hy;
        T o_yx = cur.da_yx * cur.m_yx + hy;
        T o_zx = cur.da_zx * cur.m_zx + hz;
        T o_zy = cur.da_zy * cur.m_zy + hz;
        P.apply(b / S, cur.pw, o_xy, o_xz, o_yz, o_yx, o_zx, o_zy);
        stg(mxy, b, o_xy); stg(mxz, b, o_xz); stg(myz, b, o_yz);
        stg(myx, b, o_yx); stg(mzx, b, o_zx); stg(mzy, b, o_zy);
        p_yx = cur.cb_yx * cur.v_yx;
        p_zx = cur.cb_zx * cur.v_zx;
        if (more) cur = nxt;
    }
}

// n >= 0: the adjoint sources of step n (E groups for k_adj_e, H groups for k_adj_h, from the
// instance descriptor's sorted groups) are added where the values are computed.
template <typename T, bool PF>
__global__ void __launch_bounds__(1024) k_adj_e(const InstDev<T>* __restrict__ insts, const MemberDev* __restrict__ mem,
                                                int nmem, int total, int n) {
    __shared__ int s_member;
    Tile t;
    const InstDev<T>* I = locate(insts, mem, nmem, total, &s_member, t);
    if (!I) return;
    const AdjFold<T> P{I->ebits, (const ProbeGroup<T>*)I->egroups, n >= 0 ? I->negroups : 0, n};
    if (I->coef_kind == 1) {
        const CoefTable<T> C{I->eidx, I->hidx, I->etab, I->htab};
        if (PF) adj_e_column_pf<T, CoefTable<T>>(I->F, C, I->Nx, I->Ny, I->Nz, t, P);
        else adj_e_column<T, CoefTable<T>>(I->F, C, I->Nx, I->Ny, I->Nz, t, P);
    } else {
        if (PF) adj_e_column_pf<T, CoefRaw<T>>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
        else adj_e_column<T, CoefRaw<T>>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
    }
}

// Totals from splits after a checkpoint restore (the forward computes them the same way).
template <typename T>
__global__ void __launch_bounds__(256) k_totals(const InstDev<T>* __restrict__ insts, const int64_t* __restrict__ blk0,
                                                int n) {
    const int k = item_of(blk0, n, blockIdx.x);
    const InstDev<T>& I = insts[k];
    const int64_t i = (int64_t)(blockIdx.x - blk0[k]) * blockDim.x + threadIdx.x;
    if (i >= (int64_t)I.Nx * I.Ny * I.Nz) return;
    const Fields<T>& F = I.F;
    F.ex[i] = F.ex_y[i] + F.ex_z[i];
    F.ey[i] = F.ey_z[i] + F.ey_x[i];
    F.ez[i] = F.ez_x[i] + F.ez_y[i];
    F.hx[i] = F.hx_y[i] + F.hx_z[i];
    F.hy[i] = F.hy_z[i] + F.hy_x[i];
    F.hz[i] = F.hz_x[i] + F.hz_y[i];
}

// Adjoint sources: g[n] of every probe of a group added to both splits of its component.  For
// these groups `series` holds the instance's dJ/dy, time-major [Nt][nprobes].
template <typename T>
__global__ void k_adj_inject(const InstDev<T>* __restrict__ insts, const ProbeGroup<T>* __restrict__ groups,
                             int ngroups, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ngroups) return;
    const ProbeGroup<T> Gp = groups[k];
    const Fields<T>& F = insts[Gp.inst].F;
    T *s1, *s2;
    switch (Gp.field) {
        case EX: s1 = F.ex_y; s2 = F.ex_z; break;
        case EY: s1 = F.ey_z; s2 = F.ey_x; break;
        case EZ: s1 = F.ez_x; s2 = F.ez_y; break;
        case HX: s1 = F.hx_y; s2 = F.hx_z; break;
        case HY: s1 = F.hy_z; s2 = F.hy_x; break;
        default: s1 = F.hz_x; s2 = F.hz_y; break;
    }
    const T* row = Gp.series + (size_t)n * Gp.nprobes;
    T sum = T(0);
    for (int j = 0; j < Gp.count; ++j) sum = sum + row[Gp.order[j]];
    s1[Gp.cell] = s1[Gp.cell] + sum;
    s2[Gp.cell] = s2[Gp.cell] + sum;
}

// Design box [b0, b0 + bn) in grid cells; its H slot holds (bn + 1) cells per axis.
struct Box {
    int32_t x0, y0, z0, nx, ny, nz;
};

// A design box of one instance: the per-step slots of forward H totals (K of them, slot_elems
// each), its dJ/dCb accumulator, and the block offsets of the save (hblk0) and gradient (gblk0)
// launches.
template <typename T>
struct BoxDev {
    Box B;
    int32_t inst, pad;
    T* slots;
    double* grad;
    int64_t slot_elems;
};

// The item (a design box) whose block range [blk0[i], blk0[i+1]) holds block b: every wavefront finds
// it on its own from 64 offsets per load (one memory round trip for up to 64 boxes).
__device__ __forceinline__ int box_of(const int64_t* __restrict__ blk0, int n, int64_t b) {
    const int lane = (int)(threadIdx.x & 63);
    int count = 0;
    for (int c = 0; c < n; c += 64) {
        const int k = c + lane;
        count += __popcll(__ballot(k < n && blk0[k] <= b));
    }
    return __builtin_amdgcn_readfirstlane(count - 1);
}

// Save the forward H totals of every box (plus one cell on each upper side) into slot `j`.  Box
// and grid indices fit 32 bits (a session holds at most 2^29 cells).  `block` is the workgroup's
// index among the save's workgroups (blockDim.x threads each; blk0 counts them per box).
template <typename T>
__device__ __forceinline__ void save_h_block(const InstDev<T>* __restrict__ insts, const BoxDev<T>* __restrict__ boxes,
                                             const int64_t* __restrict__ blk0, int n, int j, int block) {
    using G = GLOBAL_T(T);
    const int k = box_of(blk0, n, block);
    const BoxDev<T> D = boxes[k];
    const InstDev<T>& I = insts[D.inst];
    const uint32_t hy_n = D.B.ny + 1, hz_n = D.B.nz + 1;
    const uint32_t total = (uint32_t)(D.B.nx + 1) * hy_n * hz_n;
    const uint32_t q = (uint32_t)(block - blk0[k]) * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const uint32_t lz = q % hz_n, ly = q / hz_n % hy_n, lx = q / (hz_n * hy_n);
    const int x = D.B.x0 + (int)lx, y = D.B.y0 + (int)ly, z = D.B.z0 + (int)lz;
    T a = T(0), b = T(0), c = T(0);
    if (x < I.Nx && y < I.Ny && z < I.Nz) {
        const uint32_t i = ((uint32_t)x * I.Ny + y) * I.Nz + z;
        const Fields<T>& F = I.F;
        a = ((const G*)F.hx_y)[i] + ((const G*)F.hx_z)[i];
        b = ((const G*)F.hy_z)[i] + ((const G*)F.hy_x)[i];
        c = ((const G*)F.hz_x)[i] + ((const G*)F.hz_y)[i];
    }
    G* slot = (G*)D.slots + (size_t)j * D.slot_elems;
    slot[q] = a;
    slot[total + q] = b;
    slot[2 * total + q] = c;
}

template <typename T>
__global__ void __launch_bounds__(256) k_save_h(const InstDev<T>* __restrict__ insts, const BoxDev<T>* __restrict__ boxes,
                                                const int64_t* __restrict__ blk0, int n, int j, const int* base = nullptr) {
    if (base) j += *base;  // history recording inside a captured chunk: slot = step number
    save_h_block<T>(insts, boxes, blk0, n, j, (int)blockIdx.x);
}

// The E half step with the design-box H save of the same step in its launch: workgroups past the
// column tiles save H^n (the H splits, which the E pass only reads) into slot n.
template <typename T, bool PF, bool NT>
__global__ void __launch_bounds__(1024) k_update_e_save(const InstDev<T>* __restrict__ insts,
                                                        const MemberDev* __restrict__ mem, int nmem, int total,
                                                        const int* base, int local, int io,
                                                        const BoxDev<T>* __restrict__ boxes,
                                                        const int64_t* __restrict__ blk0, int nb) {
    {
        const int L = column_block(mem);  // the same numbering as locate's
        if (L >= total) {
            if (L - total < (int)blk0[nb]) save_h_block<T>(insts, boxes, blk0, nb, step_of(base, local), L - total);
            return;
        }
    }
    __shared__ int s_member;
    Tile t;
    const InstDev<T>* I = locate(insts, mem, nmem, total, &s_member, t);
    if (!I) return;
    const int ng = io ? I->negroups : 0;
    const ProbeFold<T> P{I->ebits, (const ProbeGroup<T>*)I->egroups, ng, ng ? step_of(base, local) : 0};
    if (I->coef_kind == 1)
        e_column<T, CoefTable<T>, PF, NT>(I->F, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, I->Nx, I->Ny, I->Nz, t, P);
    else
        e_column<T, CoefRaw<T>, PF, NT>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
}

// dJ/dCb of the six E splits in every box: += nu * (the curl difference the forward update used),
// with the forward H of this step from slot `j`.  `block` is the workgroup's index among the
// gradient's workgroups (blockDim.x threads each; blk0 counts them per box).
template <typename T>
__device__ __forceinline__ void adj_grad_block(const InstDev<T>* __restrict__ insts, const BoxDev<T>* __restrict__ boxes,
                                               const int64_t* __restrict__ blk0, int n, int j, int block) {
    using G = GLOBAL_T(T);
    using GD = GLOBAL_T(double);
    const int k = box_of(blk0, n, block);
    const BoxDev<T> D = boxes[k];
    const InstDev<T>& A = insts[D.inst];
    const Box B = D.B;
    const uint32_t hy_n = B.ny + 1, hz_n = B.nz + 1;
    const uint32_t total_h = (uint32_t)(B.nx + 1) * hy_n * hz_n;
    const uint32_t nbox = (uint32_t)B.nx * B.ny * B.nz;
    const uint32_t q = (uint32_t)(block - blk0[k]) * blockDim.x + threadIdx.x;
    if (q >= nbox) return;
    const G* hx = (const G*)D.slots + (size_t)j * D.slot_elems;
    const G* hy = hx + total_h;
    const G* hz = hx + 2 * total_h;
    const uint32_t lz = q % B.nz, ly = q / B.nz % B.ny, lx = q / ((uint32_t)B.nz * B.ny);
    const int x = B.x0 + (int)lx, y = B.y0 + (int)ly, z = B.z0 + (int)lz;
    const uint32_t i = ((uint32_t)x * A.Ny + y) * A.Nz + z;
    const uint32_t h = (lx * hy_n + ly) * hz_n + lz, hyp = h + hz_n, hxp = h + hy_n * hz_n;
    const bool xin = x < A.Nx - 1, yin = y < A.Ny - 1, zin = z < A.Nz - 1;
    const T d_xy = yin ? hz[hyp] - hz[h] : T(0);
    const T d_xz = zin ? hy[h + 1] - hy[h] : T(0);
    const T d_yz = zin ? hx[h + 1] - hx[h] : T(0);
    const T d_yx = xin ? hz[hxp] - hz[h] : T(0);
    const T d_zx = xin ? hy[hxp] - hy[h] : T(0);
    const T d_zy = yin ? hx[hyp] - hx[h] : T(0);
    const Fields<T>& F = A.F;
    GD* g = (GD*)D.grad + q;
    g[0 * (size_t)nbox] += (double)((const G*)F.ex_y)[i] * (double)d_xy;
    g[1 * (size_t)nbox] += (double)((const G*)F.ex_z)[i] * (double)d_xz;
    g[2 * (size_t)nbox] += (double)((const G*)F.ey_z)[i] * (double)d_yz;
    g[3 * (size_t)nbox] += (double)((const G*)F.ey_x)[i] * (double)d_yx;
    g[4 * (size_t)nbox] += (double)((const G*)F.ez_x)[i] * (double)d_zx;
    g[5 * (size_t)nbox] += (double)((const G*)F.ez_y)[i] * (double)d_zy;
}

template <typename T, bool PF>
__global__ void __launch_bounds__(1024) k_adj_h(const InstDev<T>* __restrict__ insts, const MemberDev* __restrict__ mem,
                                                int nmem, int total, int n, const BoxDev<T>* __restrict__ gboxes = nullptr,
                                                const int64_t* __restrict__ gblk0 = nullptr, int ngb = 0, int gj = 0,
                                                const InstDev<T>* __restrict__ ginsts = nullptr) {
    // workgroups past the column tiles (gboxes given) accumulate the design-box gradient of this
    // step in the same launch: they read the E adjoint and the saved H, which this pass does not write
    if (gboxes) {
        const int L = column_block(mem);  // the same numbering as locate's
        if (L >= total) {
            if (L - total < (int)gblk0[ngb]) adj_grad_block<T>(ginsts, gboxes, gblk0, ngb, gj, L - total);
            return;
        }
    }
    __shared__ int s_member;
    Tile t;
    const InstDev<T>* I = locate(insts, mem, nmem, total, &s_member, t);
    if (!I) return;
    const AdjFold<T> P{I->hbits, (const ProbeGroup<T>*)I->hgroups, n >= 0 ? I->nhgroups : 0, n};
    if (I->coef_kind == 1) {
        const CoefTable<T> C{I->eidx, I->hidx, I->etab, I->htab};
        if (PF) adj_h_column_pf<T, CoefTable<T>>(I->F, C, I->Nx, I->Ny, I->Nz, t, P);
        else adj_h_column<T, CoefTable<T>>(I->F, C, I->Nx, I->Ny, I->Nz, t, P);
    } else {
        if (PF) adj_h_column_pf<T, CoefRaw<T>>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
        else adj_h_column<T, CoefRaw<T>>(I->F, CoefRaw<T>{I->raw}, I->Nx, I->Ny, I->Nz, t, P);
    }
}

// ------------------------------------------------------------------------------------------------
// design-region coefficient assembly
//
// The E coefficients of a box of packed cells from cell permittivities and conductivities, as the
// host assembly forms them (rfnetwork/solver.py:892-961 averaging and Ca/Cb, :1463-1523 packing):
// the edge value of each E component is the half-width-weighted average of its 4 adjacent cells in
// float64 (the weights' sum in float32, the widths being float32; first along the lower of the two
// transverse axes, then the other), rounded to the
// element type; Ca = (2 eps - sig dt) / (2 eps + sig dt), Cb = 2 dt / (2 eps + sig dt) in that
// type; edges on the grid boundary take eps0 and no loss; the packed Cb carries the inverse dual
// spacing (zero past the last one) and the curl sign; the x = Nx-1 plane of the Ey and Ez entries
// reads zero.  Valid for edges of plain dielectric only (no PML, conductor, lumped element or edge
// correction); the check pass compares the tuple of the current permittivity with what the session
// holds.
// ------------------------------------------------------------------------------------------------

struct DesignDev {
    int32_t gx0, gy0, gz0, gnx, gny, gnz;  // packed-cell box
    int32_t ex1, ey1, ez1;                  // extents of the cell volumes: the box + 1 cell, clipped
    int32_t Nx, Ny, Nz;
    const double *eps, *sig;                // cells [gx0, gx0 + ex1) x ..., C order
    const double *dx, *dy, *dz;             // cell widths
    const void *dx1, *dy1, *dz1;            // packing inverses, element type
    double dt, e0;
};

template <typename T>
__device__ __forceinline__ void design_tuple(const DesignDev& D, int i, int j, int k, T e[12]) {
    const int li = i - D.gx0, lj = j - D.gy0, lk = k - D.gz0;
    auto cell = [&](const double* v, int a, int b, int c) {
        return v[((int64_t)(li + a) * D.ey1 + (lj + b)) * D.ez1 + (lk + c)];
    };
    // the widths are float32 values: their halves are exact in double, and the weights' sum is
    // formed in float32 as the host assembly forms it
    auto avg = [](double v0, double v1, double d0, double d1) {
        return (v0 * (d0 / 2) + v1 * (d1 / 2)) / (double)((float)(d0 / 2) + (float)(d1 / 2));
    };
    const T dt = (T)D.dt;
    auto coef = [&](bool interior, double eps64, double sig64, T& ca, T& cb) {
        const T eps = interior ? (T)eps64 : (T)D.e0;
        const T sig = interior ? (T)sig64 : T(0);
        ca = (T(2) * eps - (sig * dt)) / (T(2) * eps + (sig * dt));
        cb = (T(2) * dt) / (T(2) * eps + (sig * dt));
    };
    const double dxi = D.dx[i], dyj = D.dy[j], dzk = D.dz[k];
    const bool xi = i <= D.Nx - 2, yi = j <= D.Ny - 2, zi = k <= D.Nz - 2;
    const double dxn = xi ? D.dx[i + 1] : 0.0, dyn = yi ? D.dy[j + 1] : 0.0, dzn = zi ? D.dz[k + 1] : 0.0;
    T cax, cbx, cay, cby, caz, cbz;
    {  // Ex edge: cells (i, j..j+1, k..k+1), y then z
        double ve = 0, vs = 0;
        if (yi && zi) {
            for (int h = 0; h < 2; ++h) {
                const double* v = h ? D.sig : D.eps;
                const double u0 = avg(cell(v, 0, 0, 0), cell(v, 0, 1, 0), dyj, dyn);
                const double u1 = avg(cell(v, 0, 0, 1), cell(v, 0, 1, 1), dyj, dyn);
                (h ? vs : ve) = avg(u0, u1, dzk, dzn);
            }
        }
        coef(yi && zi, ve, vs, cax, cbx);
    }
    {  // Ey edge: cells (i..i+1, j, k..k+1), x then z
        double ve = 0, vs = 0;
        if (xi && zi) {
            for (int h = 0; h < 2; ++h) {
                const double* v = h ? D.sig : D.eps;
                const double u0 = avg(cell(v, 0, 0, 0), cell(v, 1, 0, 0), dxi, dxn);
                const double u1 = avg(cell(v, 0, 0, 1), cell(v, 1, 0, 1), dxi, dxn);
                (h ? vs : ve) = avg(u0, u1, dzk, dzn);
            }
        }
        coef(xi && zi, ve, vs, cay, cby);
    }
    {  // Ez edge: cells (i..i+1, j..j+1, k), x then y
        double ve = 0, vs = 0;
        if (xi && yi) {
            for (int h = 0; h < 2; ++h) {
                const double* v = h ? D.sig : D.eps;
                const double u0 = avg(cell(v, 0, 0, 0), cell(v, 1, 0, 0), dxi, dxn);
                const double u1 = avg(cell(v, 0, 1, 0), cell(v, 1, 1, 0), dxi, dxn);
                (h ? vs : ve) = avg(u0, u1, dyj, dyn);
            }
        }
        coef(xi && yi, ve, vs, caz, cbz);
    }
    const T ix = ((const T*)D.dx1)[i], iy = ((const T*)D.dy1)[j], iz = ((const T*)D.dz1)[k];
    e[0] = cax; e[1] = cax; e[2] = cbx * iy; e[3] = (-cbx) * iz;
    e[4] = cay; e[5] = cay; e[6] = cby * iz; e[7] = (-cby) * ix;
    e[8] = caz; e[9] = caz; e[10] = cbz * ix; e[11] = (-cbz) * iy;
    if (i == D.Nx - 1)
        for (int c = 4; c < 12; ++c) e[c] = T(0);
}

// check = 1: compare each packed cell's tuple with the session's (table entry or raw arrays) and
// mark the workgroup in bad[] on any difference; check = 0: write the tuples (table entries
// base + q and the cells' keys, or the raw arrays).
template <typename T>
__global__ void __launch_bounds__(256) k_design_coefficients(DesignDev D, int coef_kind, uint32_t* eidx,
                                                             T* etab, T* const* raw, int64_t base, int check,
                                                             int* bad) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)D.gnx * D.gny * D.gnz;
    bool diff = false;
    if (q < n) {
        const int lk = (int)(q % D.gnz), lj = (int)(q / D.gnz % D.gny), li = (int)(q / ((int64_t)D.gnz * D.gny));
        const int i = D.gx0 + li, j = D.gy0 + lj, k = D.gz0 + lk;
        const int64_t c = ((int64_t)i * D.Ny + j) * D.Nz + k;
        T e[12];
        design_tuple<T>(D, i, j, k, e);
        if (check) {
            for (int m = 0; m < 12; ++m) {
                const T have = coef_kind == 1 ? etab[(int64_t)eidx[c] * 12 + m] : raw[m][c];
                diff |= have != e[m];
            }
        } else if (coef_kind == 1) {
            for (int m = 0; m < 12; ++m) etab[(base + q) * 12 + m] = e[m];
            eidx[c] = (uint32_t)(base + q);
        } else {
            for (int m = 0; m < 12; ++m) raw[m][c] = e[m];
        }
    }
    if (check) {
        const int any = __syncthreads_or(diff ? 1 : 0);
        if (threadIdx.x == 0) bad[blockIdx.x] = any;
    }
}

// The relative permittivity of the cells of a box (C order, float or double, in device memory) into
// the held cell permittivity of a design box (F/m: eps_r * e0 in double, as the host forms it).
// (bx, by, bz): the box's first cell relative to the held volume's; (nx, ny, nz): its extents.
__global__ void __launch_bounds__(256) k_design_put_eps(double* __restrict__ eps, int ey1, int ez1, int bx, int by,
                                                        int bz, int nx, int ny, int nz, const void* __restrict__ eps_r,
                                                        int f64, double e0) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    const int k = (int)(q % nz), j = (int)(q / nz % ny), i = (int)(q / ((int64_t)nz * ny));
    const double v = f64 ? ((const double*)eps_r)[q] : (double)((const float*)eps_r)[q];
    eps[((int64_t)(bx + i) * ey1 + (by + j)) * ez1 + (bz + k)] = v * e0;
}

// ------------------------------------------------------------------------------------------------
// whole-grid coefficient assembly
//
// The host assembly (rfnetwork/solver.py:871-1003 averaging and Ca/Cb, :1239-1325 PML, :817-868
// conductors, :1086-1236 lumped elements, :2534-2700 edge corrections, :1463-1523 packing) restated
// per packed cell with the host's floating-point types: cell values and averages in float64, the
// edge permittivity rounded to the element type T, Ca/Cb of plain edges in T, the PML formulas in
// float64 (the float64 profile times eps / e0 formed in T) rounded to T at the end, H defaults
// Da = 1 and Db = dt / u0 formed in float32.  The passes: dense (unpacked values of every cell, PML
// included), the E overrides group by group, the edge-correction factors one by one in list order,
// then the packing (Cb and Db times the signed inverse spacing; the x = Nx-1 plane of arrays 4..17
// zeroed).
// ------------------------------------------------------------------------------------------------

struct AsmPml {
    int32_t axis, upper, n, pad;
    const double* edge;
    const double* half;
};

constexpr int kMaxPmlSides = 6;

template <typename T>
struct AsmDev {
    int32_t Nx, Ny, Nz, npml;
    const double* eps;
    const double* sig;
    const double* w[3];
    const T* idual[3];
    const T* iw[3];
    double dt, e0, u0;
    AsmPml pml[kMaxPmlSides];
};

__device__ __forceinline__ double asm_avg(double v0, double v1, double d0, double d1) {
    return (v0 * (d0 / 2) + v1 * (d1 / 2)) / (double)((float)(d0 / 2) + (float)(d1 / 2));
}

// permittivity (and conductivity) of E component c at packed cell (i, j, k), in T; interior edges
// average their 4 cells (Ex: y then z; Ey: x then z; Ez: x then y), boundary edges take e0 and 0
template <typename T>
__device__ __forceinline__ void asm_edge(const AsmDev<T>& A, int c, int i, int j, int k, T& eps, T& sig) {
    const int Ny = A.Ny, Nz = A.Nz;
    const bool xi = i <= A.Nx - 2, yi = j <= Ny - 2, zi = k <= Nz - 2;
    auto cell = [&](const double* v, int a, int b, int d) { return v[((int64_t)(i + a) * Ny + (j + b)) * Nz + (k + d)]; };
    bool interior;
    double ve = 0, vs = 0;
    if (c == 0) {
        interior = yi && zi;
        if (interior) {
            const double dy0 = A.w[1][j], dy1 = A.w[1][j + 1], dz0 = A.w[2][k], dz1 = A.w[2][k + 1];
            for (int h = 0; h < 2; ++h) {
                const double* v = h ? A.sig : A.eps;
                const double u0 = asm_avg(cell(v, 0, 0, 0), cell(v, 0, 1, 0), dy0, dy1);
                const double u1 = asm_avg(cell(v, 0, 0, 1), cell(v, 0, 1, 1), dy0, dy1);
                (h ? vs : ve) = asm_avg(u0, u1, dz0, dz1);
            }
        }
    } else if (c == 1) {
        interior = xi && zi;
        if (interior) {
            const double dx0 = A.w[0][i], dx1 = A.w[0][i + 1], dz0 = A.w[2][k], dz1 = A.w[2][k + 1];
            for (int h = 0; h < 2; ++h) {
                const double* v = h ? A.sig : A.eps;
                const double u0 = asm_avg(cell(v, 0, 0, 0), cell(v, 1, 0, 0), dx0, dx1);
                const double u1 = asm_avg(cell(v, 0, 0, 1), cell(v, 1, 0, 1), dx0, dx1);
                (h ? vs : ve) = asm_avg(u0, u1, dz0, dz1);
            }
        }
    } else {
        interior = xi && yi;
        if (interior) {
            const double dx0 = A.w[0][i], dx1 = A.w[0][i + 1], dy0 = A.w[1][j], dy1 = A.w[1][j + 1];
            for (int h = 0; h < 2; ++h) {
                const double* v = h ? A.sig : A.eps;
                const double u0 = asm_avg(cell(v, 0, 0, 0), cell(v, 1, 0, 0), dx0, dx1);
                const double u1 = asm_avg(cell(v, 0, 1, 0), cell(v, 1, 1, 0), dx0, dx1);
                (h ? vs : ve) = asm_avg(u0, u1, dy0, dy1);
            }
        }
    }
    eps = interior ? (T)ve : (T)A.e0;
    sig = interior ? (T)vs : T(0);
}

// permittivity of H component c (0 hx, 1 hy, 2 hz) at packed cell (i, j, k): the average of the two
// cells either side along its own axis, e0 on the boundary
template <typename T>
__device__ __forceinline__ T asm_heps(const AsmDev<T>& A, int c, int i, int j, int k) {
    const int Ny = A.Ny, Nz = A.Nz;
    const int64_t q = ((int64_t)i * Ny + j) * Nz + k;
    if (c == 0 && i <= A.Nx - 2) return (T)asm_avg(A.eps[q], A.eps[q + (int64_t)Ny * Nz], A.w[0][i], A.w[0][i + 1]);
    if (c == 1 && j <= Ny - 2) return (T)asm_avg(A.eps[q], A.eps[q + Nz], A.w[1][j], A.w[1][j + 1]);
    if (c == 2 && k <= Nz - 2) return (T)asm_avg(A.eps[q], A.eps[q + 1], A.w[2][k], A.w[2][k + 1]);
    return (T)A.e0;
}

// the PML layer (profile index) of global index g along axis a (shape L), E edges (e = true) or H
// half cells; the last side that covers it wins, as the host applies the sides in order
template <typename T>
__device__ __forceinline__ const double* asm_pml(const AsmDev<T>& A, int a, int g, int L, bool e, int& layer) {
    const double* prof = nullptr;
    for (int s = 0; s < A.npml; ++s) {
        const AsmPml& P = A.pml[s];
        if (P.axis != a) continue;
        const int n = P.n;
        if (e) {
            if (!P.upper && g >= 1 && g <= n) { layer = n - g; prof = P.edge; }
            if (P.upper && g >= L - n - 1 && g <= L - 2) { layer = g - (L - n - 1); prof = P.edge; }
        } else {
            if (!P.upper && g >= 0 && g <= n - 1) { layer = n - 1 - g; prof = P.half; }
            if (P.upper && g >= L - n && g <= L - 1) { layer = g - (L - n); prof = P.half; }
        }
    }
    return prof;
}

template <typename T>
__global__ void __launch_bounds__(256) k_asm_dense(AsmDev<T> A, T* const* raw) {
    const int64_t n = (int64_t)A.Nx * A.Ny * A.Nz;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int k = (int)(q % A.Nz), j = (int)(q / A.Nz % A.Ny), i = (int)(q / ((int64_t)A.Nz * A.Ny));
    const int p[3] = {i, j, k};
    const int N[3] = {A.Nx, A.Ny, A.Nz};
    const T dt = (T)A.dt;
    const float dtf = (float)A.dt;
    // E: split (component c, axis a) -> Ca at raw[ca], Cb at raw[ca + 2]
    const int esplit[6][3] = {{0, 1, 0}, {0, 2, 1}, {1, 2, 4}, {1, 0, 5}, {2, 0, 8}, {2, 1, 9}};
    T eps[3], sig[3];
    for (int c = 0; c < 3; ++c) asm_edge<T>(A, c, i, j, k, eps[c], sig[c]);
    for (int s = 0; s < 6; ++s) {
        const int c = esplit[s][0], a = esplit[s][1], ca = esplit[s][2];
        const T e = eps[c], sg = sig[c];
        T Ca = (T(2) * e - (sg * dt)) / (T(2) * e + (sg * dt));
        T Cb = (T)(2.0f * dtf) / (T(2) * e + (sg * dt));
        int layer = 0;
        const double* prof = asm_pml<T>(A, a, p[a] + 1, N[a] + 1, true, layer);
        if (prof) {
            const double sp = prof[layer] * (double)(e / (T)A.e0);
            const double sdt = sp * (double)dtf;
            const double te = (double)(T(2) * e);
            Ca = (T)((te - sdt) / (te + sdt));
            Cb = (T)((double)(2.0f * dtf) / (te + sdt));
        }
        raw[ca][q] = Ca;
        raw[ca + 2][q] = Cb;
    }
    // H: split (component c, axis a) -> Da at raw[da], Db1 and Db2 at raw[db], raw[db + 1]
    const int hsplit[6][4] = {{0, 1, 12, 14}, {0, 2, 13, 16}, {1, 2, 18, 20}, {1, 0, 19, 22}, {2, 0, 24, 26}, {2, 1, 25, 28}};
    const T db0 = (T)(dtf / (float)A.u0);
    for (int s = 0; s < 6; ++s) {
        const int c = hsplit[s][0], a = hsplit[s][1], da = hsplit[s][2], db = hsplit[s][3];
        T Da = T(1), Db = db0;
        int layer = 0;
        const double* prof = asm_pml<T>(A, a, p[a], N[a], false, layer);
        if (prof) {
            const T e = asm_heps<T>(A, c, i, j, k);
            const double sp = prof[layer] * (double)(e / (T)A.e0);
            const double sm = sp * A.u0 / (double)e;
            const double smdt = sm * (double)dtf;
            Da = (T)((2 * A.u0 - smdt) / (2 * A.u0 + smdt));
            Db = (T)((double)(2.0f * dtf) / (2 * A.u0 + smdt));
        }
        raw[da][q] = Da;
        raw[db][q] = Db;
        raw[db + 1][q] = Db;
    }
}

// one override group: Ca and Cb of both splits of component c at the listed packed cells
template <typename T>
__global__ void __launch_bounds__(256) k_asm_fix(AsmDev<T> A, T* const* raw, int c, int kind, int wide,
                                                 int64_t count, const int64_t* cells, const double* values) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t q = cells[t];
    const int k = (int)(q % A.Nz), j = (int)(q / A.Nz % A.Ny), i = (int)(q / ((int64_t)A.Nz * A.Ny));
    T e, sg;
    asm_edge<T>(A, c, i, j, k, e, sg);
    const T dt = (T)A.dt;
    const float dtf = (float)A.dt;
    T Ca, Cb;
    if (kind == 0) {  // conductor: s = sigma dt
        const double s = values[0];
        if (wide) {
            const double te = (double)(T(2) * e);
            Ca = (T)((te - s) / (te + s));
            Cb = (T)((double)(2.0f * dtf) / (te + s));
        } else {
            const T st = (T)(float)s;
            Ca = (T(2) * e - st) / (T(2) * e + st);
            Cb = (T)(2.0f * dtf) / (T(2) * e + st);
        }
    } else {  // lumped element: q = df / (2 r dn dw)
        const double qv = values[t];
        const T a = e / dt;
        if (wide) {
            const double den = (double)a + qv;
            Ca = (T)(((double)a - qv) / den);
            Cb = (T)(1.0 / den);
        } else {
            const T qt = (T)(float)qv;
            const T den = a + qt;
            Ca = (a - qt) / den;
            Cb = T(1) / den;
        }
    }
    const int ca = 4 * c;  // Ca of the two splits at 4c, 4c + 1; Cb at 4c + 2, 4c + 3
    raw[ca][q] = Ca;
    raw[ca + 1][q] = Ca;
    raw[ca + 2][q] = Cb;
    raw[ca + 3][q] = Cb;
}

// edge-correction factors, one thread in list order (an entry may be scaled more than once)
template <typename T>
__global__ void k_asm_scale(T* const* raw, int n, const rfdtd_h_scale* sc) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int m = 0; m < n; ++m) {
        T* v = raw[sc[m].coef] + sc[m].cell;
        if (sc[m].wide || sizeof(T) == sizeof(double))
            *v = (T)((double)*v * sc[m].factor);
        else
            *v = (T)((float)*v * (float)sc[m].factor);
    }
}

// packing: Cb and Db times the signed inverse spacing; arrays 4..17 read zero on x = Nx-1
template <typename T>
__global__ void __launch_bounds__(256) k_asm_pack(AsmDev<T> A, T* const* raw) {
    const int64_t n = (int64_t)A.Nx * A.Ny * A.Nz;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int k = (int)(q % A.Nz), j = (int)(q / A.Nz % A.Ny), i = (int)(q / ((int64_t)A.Nz * A.Ny));
    // array, sign, axis, table (0: inverse dual spacing, 1: inverse width)
    const int8_t pk[20][4] = {{2, 1, 1, 0},  {3, -1, 2, 0}, {6, 1, 2, 0},  {7, -1, 0, 0},  {10, 1, 0, 0},
                              {11, -1, 1, 0}, {14, -1, 1, 1}, {15, -1, 1, 1}, {16, 1, 2, 1},  {17, 1, 2, 1},
                              {20, -1, 2, 1}, {21, -1, 2, 1}, {22, 1, 0, 1},  {23, 1, 0, 1},  {26, -1, 0, 1},
                              {27, -1, 0, 1}, {28, 1, 1, 1},  {29, 1, 1, 1},  {0, 0, 0, 0},   {0, 0, 0, 0}};
    const int p[3] = {i, j, k};
    for (int m = 0; m < 18; ++m) {
        const int a = pk[m][2];
        const T inv = pk[m][3] ? A.iw[a][p[a]] : A.idual[a][p[a]];
        T* v = raw[pk[m][0]] + q;
        *v = (pk[m][1] < 0 ? -*v : *v) * inv;
    }
    if (i == A.Nx - 1)
        for (int m = 4; m <= 17; ++m) raw[m][q] = T(0);
}

// The monitor whose block range holds this workgroup (blk0: sampling launch, fblk0: flush launch):
// the last one whose first block is at most blockIdx.x.  Every wavefront finds it on its own from 64
// table entries per load, so the lookup costs one memory round trip (no serial walk, no barrier).
__device__ __forceinline__ int monitor_of(const MonitorDev* __restrict__ mons, int nmon, int block, bool flush = false) {
    const int lane = (int)(threadIdx.x & 63);
    int count = 0;
    for (int b = 0; b < nmon; b += 64) {
        const int k = b + lane;
        const int first = k < nmon ? (flush ? mons[k].fblk0 : mons[k].blk0) : INT_MAX;
        count += __popcll(__ballot(first <= block));
    }
    return __builtin_amdgcn_readfirstlane(count - 1);
}

// Sample step n: time-domain monitors write their plane row; phasor monitors copy the plane into
// staging slot `slot` and record the phase row of the sample (-1: not sampled at this step).
// totals = 1 when the step's kernels stored the six totals (the separate E and H kernels do): the
// sample is then one load of the total instead of two of its splits (the same sum, formed once).
template <typename T>
__global__ void k_monitors(const InstDev<T>* __restrict__ insts, const MonitorDev* __restrict__ mons, int nmon,
                           const int* base, int local, int slot, int* __restrict__ rows, int totals) {
    const MonitorDev M = mons[monitor_of(mons, nmon, (int)blockIdx.x)];
    const int n = step_of(base, local);
    const bool sampled = n == 0 || (n % M.n_step) == 0;
    const int64_t q = (int64_t)(blockIdx.x - M.blk0) * blockDim.x + threadIdx.x;
    if (M.nf && q == 0) rows[M.id * kStageSlots + slot] = sampled ? n / M.n_step : -1;
    if (!sampled || q >= M.N1N2) return;
    const InstDev<T>& I = insts[M.inst];
    const int Ny = I.Ny, Nz = I.Nz;
    const int a = (int)(q / M.N2), b = (int)(q - (int64_t)a * M.N2);
    int64_t idx;
    if (M.axis == 0)
        idx = ((int64_t)M.pos * Ny + a) * Nz + b;
    else if (M.axis == 1)
        idx = ((int64_t)a * Ny + M.pos) * Nz + b;
    else
        idx = ((int64_t)a * Ny + b) * Nz + M.pos;
    const T* tot[6] = {I.F.ex, I.F.ey, I.F.ez, I.F.hx, I.F.hy, I.F.hz};
    const T v = totals ? ((const GLOBAL_T(T)*)tot[M.field])[idx] : total_at(I, M.field, idx);
    if (M.nf == 0) {
        const int64_t m = n / M.n_step;
        ((GLOBAL_T(T)*)M.values)[(M.ring ? m % M.ring : m) * M.N1N2 + q] = v;
    } else {
        ((GLOBAL_T(T)*)M.stage)[(int64_t)slot * M.N1N2 + q] = v;
    }
}

// Frequencies per flush thread: their accumulators are loaded together.
template <typename A>
constexpr int flush_group() { return sizeof(A) == sizeof(double) ? kFreqGroup / 2 : kFreqGroup; }

// Accumulate staged slots [0, nslots): values[f] += v_s * phase[row_s, f] for s in time order (a real
// field times a complex phase: one product and one sum per component, as each step would).  A thread
// owns one plane cell and one group of frequencies; the groups of a block of cells are consecutive
// logical workgroups, which xcd_logical keeps on one XCD, so the staged samples that every group
// reads come from that XCD's L2 after the first.  Within a group the samples are taken in time
// order, each against the group's contiguous run of phase entries.  A is the accumulator and phase
// type: T, or double for a float session's monitor accumulated in double (each float sample is
// widened exactly, then multiplied and added in double).
template <typename T, typename A>
__device__ __forceinline__ void flush_monitor(const MonitorDev& M, int local, int nslots, const int* __restrict__ rows) {
    constexpr int G = flush_group<A>();
    const int ngroups = (M.nf + G - 1) / G;
    const int cblock = local / ngroups;
    const int f0 = (local - cblock * ngroups) * G;
    const int64_t q = (int64_t)cblock * blockDim.x + threadIdx.x;
    if (q >= M.N1N2) return;
    const int nf = M.nf;
    const int64_t fs = 2 * M.N1N2;  // between the accumulators of consecutive frequencies
    GLOBAL_T(A)* acc = (GLOBAL_T(A)*)M.values + 2 * q + f0 * fs;
    A re[G], im[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        re[g] = f0 + g < nf ? acc[g * fs] : A(0);
        im[g] = f0 + g < nf ? acc[g * fs + 1] : A(0);
    }
    A v[kStageSlots];
    int r[kStageSlots];
    // the sample rows and the phase table are read-only during the flush and the same for every
    // lane: scalar loads
    const CONST_AS int* srows = (const CONST_AS int*)rows + M.id * kStageSlots;
    const GLOBAL_T(T)* stage = (const GLOBAL_T(T)*)M.stage + q;
#pragma unroll
    for (int j = 0; j < kStageSlots; ++j) {
        r[j] = j < nslots ? srows[j] : -1;
        v[j] = r[j] >= 0 ? (A)stage[(int64_t)j * M.N1N2] : A(0);
    }
    const CONST_AS A* phase = (const CONST_AS A*)M.phase;
#pragma unroll
    for (int j = 0; j < kStageSlots; ++j) {
        if (r[j] < 0) continue;
        const CONST_AS A* ph = phase + ((size_t)r[j] * nf + f0) * 2;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (f0 + g >= nf) break;
            re[g] = re[g] + v[j] * ph[2 * g];
            im[g] = im[g] + v[j] * ph[2 * g + 1];
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (f0 + g < nf) {
            acc[g * fs] = re[g];
            acc[g * fs + 1] = im[g];
        }
    }
}

// Launched on a multiple of 8 workgroups (xcd_logical); logical blocks past `total` have no work.
template <typename T>
__global__ void k_monitor_flush(const MonitorDev* __restrict__ mons, int nmon, int total, int nslots,
                                const int* __restrict__ rows) {
    const int L = xcd_logical(blockIdx.x, gridDim.x);
    if (L >= total) return;
    const MonitorDev M = mons[monitor_of(mons, nmon, L, true)];
    if (M.nf == 0) return;
    if constexpr (std::is_same<T, float>::value) {
        if (M.acc64) {
            flush_monitor<float, double>(M, L - M.fblk0, nslots, rows);
            return;
        }
    }
    flush_monitor<T, T>(M, L - M.fblk0, nslots, rows);
}

// ------------------------------------------------------------------------------------------------
// two-step pass: time steps n and n + 1 in one pass over the fields (temporal blocking)
//
// A workgroup of TY + 4 wavefronts owns a tile of TY rows (y) x kP2Fz cells (z) and a range
// [x0, x1) of x planes; lane l of wave w sits at y = y0 - 2 + w, z = z0 - 2 + l.  Each step shrinks
// the region whose values are right by one cell on every side:
//   step n:      E on w <= TY + 2, l <= 62;        H on 1 <= w, l and the same upper limits;
//   step n + 1:  E on 1 <= w <= TY + 1, 1 <= l <= 61;  H (owned, stored) on 2 <= w, l likewise.
// In x the pass is skewed by one plane: iteration i takes plane i through step n and plane i - 1
// through step n + 1, the x neighbours carried in registers, the y/z neighbours exchanged through
// LDS (three barriers per plane).  Step n runs on planes x0 - 2 .. x1: the planes outside [x0, x1)
// and the halo lanes compute values the tile needs but does not own.  Only step n + 1's splits are
// stored, into the other buffer of the instance's pair (a tile's halo reads must see the old state);
// step n's monitor samples are written here by the lanes that own the cells, step n + 1's by
// k_monitors from the stored splits.  Sources are applied by every lane that computes a cell (halo
// copies included; the samples come from the pristine series), records are written by the owner.
// With two == 0 the pass advances step n alone and stores it (planes x0 - 1 .. x1 - 1).  Every value
// is formed by the same operations in the same order as in the separate kernels.
// ------------------------------------------------------------------------------------------------

// Rows owned by a two-step tile: 12 (16 wavefronts, one workgroup per CU at the pass's register
// count) or 4 (8 wavefronts: two workgroups per CU, each hiding the other's barriers and loads, for
// 50% more halo reads).
constexpr int kP2TyTall = 12, kP2TyShort = 4;
constexpr int kP2TySkew = 8;  // the one-barrier pass (more registers: 12 wavefronts, one workgroup per CU)
constexpr int kP2Fz = 60;   // z cells owned by a two-step tile of 64-lane rows
// Tiles of 32-lane rows (two rows per wavefront; 28 z cells owned per row): 28 rows in 1024 threads
// (76.5% of the lanes own a cell, against 70.3% for 12 x 60 in 64-lane rows), or 12 rows in 512
// threads (two workgroups per CU); narrow grids (Nz = 64) waste less of their second z tile.
constexpr int kP2Ty32Tall = 28, kP2Ty32Short = 12, kP2Ty32Mid = 26;
// cell-updates per step (cells x excitations, all grids) from which a session starts on the two-step pass
constexpr int64_t kTwoStepDefaultWork = (int64_t)3 << 19;  // 1.5 Mi
constexpr int kP2Mon = 64;  // monitors a two-step pass samples itself, at most

// The monitor samples of one owned cell at step n: every monitor of the tile's list (those of this
// instance that sample step n and whose plane crosses the tile) whose plane holds the cell and whose
// field is component fbase + 0..2 takes v[component] (k_monitors' placement).
template <typename T>
__device__ __forceinline__ void p2_sample(const MonitorDev* __restrict__ mons, const int* list, int nlist, int fbase,
                                          int x, int y, int z, int n, int slot, T vx, T vy, T vz) {
    for (int k = 0; k < nlist; ++k) {
        const MonitorDev& M = mons[__builtin_amdgcn_readfirstlane(list[k])];
        const int f = M.field - fbase;
        if (f < 0 || f > 2) continue;
        int a, b;
        if (M.axis == 0) {
            if (x != M.pos) continue;
            a = y; b = z;
        } else if (M.axis == 1) {
            if (y != M.pos) continue;
            a = x; b = z;
        } else {
            if (z != M.pos) continue;
            a = x; b = y;
        }
        const int64_t q = (int64_t)a * M.N2 + b;
        const T v = f == 0 ? vx : f == 1 ? vy : vz;
        if (M.nf == 0) {
            const int64_t m = n / M.n_step;
            ((GLOBAL_T(T)*)M.values)[(M.ring ? m % M.ring : m) * M.N1N2 + q] = v;
        } else {
            ((GLOBAL_T(T)*)M.stage)[(int64_t)slot * M.N1N2 + q] = v;
        }
    }
}

template <typename T, int TY, class CA, int RS = 64>
__device__ __forceinline__ void two_step_tile(const InstDev<T>& I, const CA& C, int x0, int x1, int y0, int z0, int n,
                                              bool io, bool two, const MonitorDev* __restrict__ mons,
                                              const int* mlist, int nmon, int slot, T* sm) {
    using G = GLOBAL_T(T);
    using G32 = GLOBAL_AS uint32_t;
    constexpr int NT = (TY + 4) * RS;  // lanes: TY + 4 rows of RS lanes (64, or 32: two rows per wavefront)
    constexpr uint32_t S = (uint32_t)sizeof(T);
    constexpr bool kEarly = std::is_same<CA, CoefTable<T>>::value;  // all gathers at the iteration's top
    const int Nx = I.Nx, Ny = I.Ny, Nz = I.Nz;
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S;
    const int tid = threadIdx.x, w = tid / RS, l = tid % RS;  // row and lane in the row
    const int y = y0 - 2 + w, z = z0 - 2 + l;
    const bool valid = y >= 0 && y < Ny && z >= 0 && z < Nz;
    const bool mAe = valid && w <= TY + 2 && l <= RS - 2;  // E of step n right here
    const bool mAh = mAe && w >= 1 && l >= 1;              // H of step n
    const bool mBe = mAh && w <= TY + 1 && l <= RS - 3;    // E of step n + 1
    const bool own = mBe && w >= 2 && l >= 2;            // H of step n + 1; stored, recorded
    const bool yin = y < Ny - 1, zin = z < Nz - 1;
    T* const s0 = sm;              // [3][NT] input H totals, plane i
    T* const s1e = sm + 3 * NT;    // [3][NT] E totals after step n, plane i
    T* const s1h = sm + 6 * NT;    // [2][3][NT] H totals after step n, by plane parity
    T* const s2e = sm + 12 * NT;   // [3][NT] E totals after step n + 1, plane i - 1
    const Fields<T>& A = I.F;
    const Fields<T>& B = I.G;
    const int a0 = two ? max(x0 - 2, 0) : max(x0 - 1, 0);  // first plane through step n
    const int a1 = two ? min(x1 + 1, Nx) : x1;             // one past the last
    const int b0 = max(x0 - 1, 0);                         // first plane through step n + 1 (E)
    const int ie = two ? x1 + 1 : x1;                      // iterations i in [a0, ie)
    const uint32_t p = valid ? ((uint32_t)y * (uint32_t)Nz + (uint32_t)z) : 0u;
    const uint32_t base = ((uint32_t)a0 * (uint32_t)Ny * (uint32_t)Nz + p) * S;
    auto off = [&](int x) { return base + (uint32_t)(x - a0) * sx; };
    auto load_h = [&](int x) {
        Split6<T> h{};
        if (valid && x < Nx) {
            const uint32_t b = off(x);
            h.xy = ldg((const G*)A.hx_y, b); h.xz = ldg((const G*)A.hx_z, b); h.yz = ldg((const G*)A.hy_z, b);
            h.yx = ldg((const G*)A.hy_x, b); h.zx = ldg((const G*)A.hz_x, b); h.zy = ldg((const G*)A.hz_y, b);
        }
        return h;
    };
    auto load_e = [&](int x) {
        Split6<T> e{};
        if (mAe && x < a1) {
            const uint32_t b = off(x);
            e.xy = ldg((const G*)A.ex_y, b); e.xz = ldg((const G*)A.ex_z, b); e.yz = ldg((const G*)A.ey_z, b);
            e.yx = ldg((const G*)A.ey_x, b); e.zx = ldg((const G*)A.ez_x, b); e.zy = ldg((const G*)A.ez_y, b);
        }
        return e;
    };
    auto store = [&](const Fields<T>& F, uint32_t b, const Split6<T>& e, bool h) {
        if (h) {
            stg((G*)F.hx_y, b, e.xy); stg((G*)F.hx_z, b, e.xz); stg((G*)F.hy_z, b, e.yz);
            stg((G*)F.hy_x, b, e.yx); stg((G*)F.hz_x, b, e.zx); stg((G*)F.hz_y, b, e.zy);
        } else {
            stg((G*)F.ex_y, b, e.xy); stg((G*)F.ex_z, b, e.xz); stg((G*)F.ey_z, b, e.yz);
            stg((G*)F.ey_x, b, e.yx); stg((G*)F.ez_x, b, e.zx); stg((G*)F.ez_y, b, e.zy);
        }
    };
    const bool eprobes = io && I.negroups, hprobes = io && I.nhgroups;
    // cell index of a valid lane's cell on plane x (grids hold under 2^32 bytes per array: 32 bits)
    auto cell_of = [&](int x) { return ((uint32_t)x * (uint32_t)Ny + (uint32_t)y) * (uint32_t)Nz + (uint32_t)z; };
    // probe bitmask words of a plane's cell, loaded with the plane's inputs (0 without probes)
    auto ewd = [&](int x) { return eprobes && mAe && x < a1 ? ldg((const G32*)I.ebits, (cell_of(x) >> 5) * 4u) : 0u; };
    auto hwd = [&](int x) { return hprobes && mAh && x < a1 ? ldg((const G32*)I.hbits, (cell_of(x) >> 5) * 4u) : 0u; };
    auto bit = [&](uint32_t word, int x) { return ((word >> (cell_of(x) & 31u)) & 1u) != 0; };

    Split6<T> h0 = load_h(a0), e0 = load_e(a0);  // input H and E splits of plane i
    Split6<T> hN = load_h(a0 + 1);                // input H splits of plane i + 1
    uint32_t ek = mAe ? C.ekey(off(a0)) : 0u, hk = mAh ? C.hkey(off(a0)) : 0u;
    uint32_t ew = ewd(a0), hw = hwd(a0);         // bitmask words of plane i
    Split6<T> e1p{}, h1p{};                      // splits after step n, plane i - 1
    uint32_t ekp = 0u, hkp = 0u, ewp = 0u, hwp = 0u;  // keys and bitmask words of plane i - 1
    T e1y_m = T(0), e1z_m = T(0);                // Ey, Ez after step n, plane i - 1
    T e2y_m = T(0), e2z_m = T(0);                // Ey, Ez after step n + 1, plane i - 2
    for (int i = a0; i < ie; ++i) {
        const bool doA = i < a1;  // uniform over the workgroup, as is every branch around a barrier
        const int j = i - 1;
        const bool doB = two && j >= b0 && j < x1;
        const bool recA = own && i >= x0 && i < x1;
        const bool recB = doB && own && j >= x0;
        // Table form: every coefficient gather of this iteration first (the keys came an iteration
        // ago), step n's E and H coefficients of plane i and step n + 1's of plane i - 1; the next
        // planes' inputs are issued after them, once step n's E is done.  Loads complete in issue
        // order, so a gather issued after those loads could only be waited for together with them;
        // in this order they stay in flight into the next iteration while this plane's phases wait
        // for the gathers.  Raw arrays (30 loads per cell instead of 8 table rows) would need more
        // registers than the 16-wavefront workgroup allows: there each phase's coefficients are
        // issued a phase early.
        ECoef<T> cA{};
        if (doA && mAe) cA = C.e_at(ek);
        HCoef<T> chA{};
        ECoef<T> cB{};
        HCoef<T> chB{};
        if constexpr (kEarly) {
            if (doA && mAh) chA = C.h_at(hk);
            if (doB && mBe) cB = C.e_at(ekp);
            if (recB) chB = C.h_at(hkp);
        }
        // ---- step n, plane i: E
        const T h0x = h0.xy + h0.xz, h0y = h0.yz + h0.yx, h0z = h0.zx + h0.zy;
        s0[tid] = h0x;
        s0[NT + tid] = h0y;
        s0[2 * NT + tid] = h0z;
        __syncthreads();
        if constexpr (!kEarly) {
            if (doA && mAh) chA = C.h_at(hk);
        }
        Split6<T> e1 = e0;
        T e1x = T(0), e1y = T(0), e1z = T(0);
        if (doA && mAe) {
            const ECoef<T>& c = cA;
            const bool xin = i < Nx - 1;
            const T hz_yn = yin ? s0[2 * NT + tid + RS] : T(0), hx_yn = yin ? s0[tid + RS] : T(0);
            const T hy_zn = zin ? s0[NT + tid + 1] : T(0), hx_zn = zin ? s0[tid + 1] : T(0);
            const T hy_xn = xin ? hN.yz + hN.yx : T(0), hz_xn = xin ? hN.zx + hN.zy : T(0);
            if (yin) e1.xy = c.ca_xy * e1.xy + c.cb_xy * (hz_yn - h0z);
            if (zin) e1.xz = c.ca_xz * e1.xz + c.cb_xz * (hy_zn - h0y);
            if (zin) e1.yz = c.ca_yz * e1.yz + c.cb_yz * (hx_zn - h0x);
            if (xin) e1.yx = c.ca_yx * e1.yx + c.cb_yx * (hz_xn - h0z);
            if (xin) e1.zx = c.ca_zx * e1.zx + c.cb_zx * (hy_xn - h0y);
            if (yin) e1.zy = c.ca_zy * e1.zy + c.cb_zy * (hx_yn - h0x);
            if (eprobes && bit(ew, i))
                inject_e(I, cell_of(i), n, recA, e1.xy, e1.xz, e1.yz, e1.yx, e1.zx, e1.zy);
            e1x = e1.xy + e1.xz;
            e1y = e1.yz + e1.yx;
            e1z = e1.zx + e1.zy;
            if (recA) {
                if (!two) store(B, off(i), e1, false);
                else if (nmon) p2_sample(mons, mlist, nmon, EX, i, y, z, n, slot, e1x, e1y, e1z);
            }
        }
        if constexpr (!kEarly) {
            if (doB && mBe) cB = C.e_at(ekp);
        }
        // the inputs of the next planes, in flight until the next iteration: H splits two planes
        // ahead (step n's E of plane i + 1 reads plane i + 2), E splits, keys and bitmask words one
        Split6<T> hNN{};  // read by the next iteration's E of plane i + 1 while that plane is in [a0, a1)
        if (i + 2 <= a1) hNN = load_h(i + 2);
        Split6<T> eN{};
        uint32_t ekN = 0u, hkN = 0u;
        if (i + 1 < a1) {
            eN = load_e(i + 1);
            if (mAe) ekN = C.ekey(off(i + 1));
            if (mAh) hkN = C.hkey(off(i + 1));
        }
        const uint32_t ewN = ewd(i + 1), hwN = hwd(i + 1);
        s1e[tid] = e1x;
        s1e[NT + tid] = e1y;
        s1e[2 * NT + tid] = e1z;
        __syncthreads();
        // ---- step n, plane i: H
        Split6<T> h1 = h0;
        T h1x = T(0), h1y = T(0), h1z = T(0);
        if (doA && mAh) {
            const HCoef<T>& ch = chA;
            const T ez_ym = y > 0 ? s1e[2 * NT + tid - RS] : T(0), ex_ym = y > 0 ? s1e[tid - RS] : T(0);
            const T ey_zm = z > 0 ? s1e[NT + tid - 1] : T(0), ex_zm = z > 0 ? s1e[tid - 1] : T(0);
            h1.xy = ch.da_xy * h0.xy + (ch.db_xy2 * e1z - ch.db_xy1 * ez_ym);
            h1.xz = ch.da_xz * h0.xz + (ch.db_xz2 * e1y - ch.db_xz1 * ey_zm);
            h1.yz = ch.da_yz * h0.yz + (ch.db_yz2 * e1x - ch.db_yz1 * ex_zm);
            h1.yx = ch.da_yx * h0.yx + (ch.db_yx2 * e1z - ch.db_yx1 * e1z_m);
            h1.zx = ch.da_zx * h0.zx + (ch.db_zx2 * e1y - ch.db_zx1 * e1y_m);
            h1.zy = ch.da_zy * h0.zy + (ch.db_zy2 * e1x - ch.db_zy1 * ex_ym);
            if (hprobes && bit(hw, i))
                inject_cell((const ProbeGroup<T>*)I.hgroups, I.nhgroups, cell_of(i), n, recA, h1.xy, h1.xz, h1.yz,
                            h1.yx, h1.zx, h1.zy);
            h1x = h1.xy + h1.xz;
            h1y = h1.yz + h1.yx;
            h1z = h1.zx + h1.zy;
            if (recA) {
                if (!two) store(B, off(i), h1, true);
                else if (nmon) p2_sample(mons, mlist, nmon, HX, i, y, z, n, slot, h1x, h1y, h1z);
            }
        }
        if constexpr (!kEarly) {
            if (recB) chB = C.h_at(hkp);
        }
        if (two) {
            T* const sh = s1h + (i & 1) * 3 * NT;
            sh[tid] = h1x;
            sh[NT + tid] = h1y;
            sh[2 * NT + tid] = h1z;
        }
        e1y_m = e1y;
        e1z_m = e1z;
        // ---- step n + 1, plane j = i - 1 (its H-after-step-n totals were exchanged an iteration ago)
        if (doB) {
            const T* const shj = s1h + (j & 1) * 3 * NT;
            Split6<T> e2 = e1p;
            T e2x = T(0), e2y = T(0), e2z = T(0);
            if (mBe) {
                const ECoef<T>& c = cB;
                const bool xin = j < Nx - 1;
                const T hx_c = h1p.xy + h1p.xz, hy_c = h1p.yz + h1p.yx, hz_c = h1p.zx + h1p.zy;
                const T hz_yn = yin ? shj[2 * NT + tid + RS] : T(0), hx_yn = yin ? shj[tid + RS] : T(0);
                const T hy_zn = zin ? shj[NT + tid + 1] : T(0), hx_zn = zin ? shj[tid + 1] : T(0);
                const T hy_xn = xin ? h1y : T(0), hz_xn = xin ? h1z : T(0);  // plane i, this iteration's step n
                if (yin) e2.xy = c.ca_xy * e2.xy + c.cb_xy * (hz_yn - hz_c);
                if (zin) e2.xz = c.ca_xz * e2.xz + c.cb_xz * (hy_zn - hy_c);
                if (zin) e2.yz = c.ca_yz * e2.yz + c.cb_yz * (hx_zn - hx_c);
                if (xin) e2.yx = c.ca_yx * e2.yx + c.cb_yx * (hz_xn - hz_c);
                if (xin) e2.zx = c.ca_zx * e2.zx + c.cb_zx * (hy_xn - hy_c);
                if (yin) e2.zy = c.ca_zy * e2.zy + c.cb_zy * (hx_yn - hx_c);
                if (eprobes && bit(ewp, j))
                    inject_e(I, cell_of(j), n + 1, recB, e2.xy, e2.xz, e2.yz, e2.yx, e2.zx, e2.zy);
                e2x = e2.xy + e2.xz;
                e2y = e2.yz + e2.yx;
                e2z = e2.zx + e2.zy;
                if (recB) store(B, off(j), e2, false);
            }
            s2e[tid] = e2x;
            s2e[NT + tid] = e2y;
            s2e[2 * NT + tid] = e2z;
            __syncthreads();
            if (recB) {
                const HCoef<T>& c = chB;
                const T ez_ym = y > 0 ? s2e[2 * NT + tid - RS] : T(0), ex_ym = y > 0 ? s2e[tid - RS] : T(0);
                const T ey_zm = z > 0 ? s2e[NT + tid - 1] : T(0), ex_zm = z > 0 ? s2e[tid - 1] : T(0);
                Split6<T> h2;
                h2.xy = c.da_xy * h1p.xy + (c.db_xy2 * e2z - c.db_xy1 * ez_ym);
                h2.xz = c.da_xz * h1p.xz + (c.db_xz2 * e2y - c.db_xz1 * ey_zm);
                h2.yz = c.da_yz * h1p.yz + (c.db_yz2 * e2x - c.db_yz1 * ex_zm);
                h2.yx = c.da_yx * h1p.yx + (c.db_yx2 * e2z - c.db_yx1 * e2z_m);
                h2.zx = c.da_zx * h1p.zx + (c.db_zx2 * e2y - c.db_zx1 * e2y_m);
                h2.zy = c.da_zy * h1p.zy + (c.db_zy2 * e2x - c.db_zy1 * ex_ym);
                if (hprobes && bit(hwp, j))
                    inject_cell((const ProbeGroup<T>*)I.hgroups, I.nhgroups, cell_of(j), n + 1, true, h2.xy, h2.xz,
                                h2.yz, h2.yx, h2.zx, h2.zy);
                store(B, off(j), h2, true);
            }
            e2y_m = e2y;
            e2z_m = e2z;
        }
        e1p = e1;
        h1p = h1;
        ekp = ek;
        hkp = hk;
        ewp = ew;
        hwp = hw;
        h0 = hN;
        hN = hNN;
        e0 = eN;
        ek = ekN;
        hk = hkN;
        ew = ewN;
        hw = hwN;
    }
}

// The two-step pass with one barrier per plane (SKEW): iteration i advances E of step n on plane i,
// H of step n on plane i - 1, E of step n + 1 on plane i - 2 and H of step n + 1 on plane i - 3.  Each
// y/z neighbour a phase reads was put in LDS an iteration earlier (the input H totals: before this
// iteration's barrier), every exchange alternating between two buffers by plane parity, so the one
// barrier orders them all, and the four phases' coefficient gathers are in flight across it.  The
// x neighbours and the states the later phases advance pass in registers (planes i - 1 .. i - 3);
// the input H splits are loaded two planes ahead.  Tiles, halos, sources, records and monitor
// samples as in two_step_tile, whose single-step mode serves the odd steps at the same tiling.
template <typename T, int TY, class CA>
__device__ __forceinline__ void two_step_skew_tile(const InstDev<T>& I, const CA& C, int x0, int x1, int y0, int z0,
                                                   int n, bool io, const MonitorDev* __restrict__ mons,
                                                   const int* mlist, int nmon, int slot, T* sm) {
    using G = GLOBAL_T(T);
    using G32 = GLOBAL_AS uint32_t;
    constexpr int NT = (TY + 4) * 64;
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const int Nx = I.Nx, Ny = I.Ny, Nz = I.Nz;
    const uint32_t sx = (uint32_t)Ny * (uint32_t)Nz * S;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int y = y0 - 2 + w, z = z0 - 2 + l;
    const bool valid = y >= 0 && y < Ny && z >= 0 && z < Nz;
    const bool mAe = valid && w <= TY + 2 && l <= 62;  // E of step n right here
    const bool mAh = mAe && w >= 1 && l >= 1;          // H of step n
    const bool mBe = mAh && w <= TY + 1 && l <= 61;    // E of step n + 1
    const bool own = mBe && w >= 2 && l >= 2;          // H of step n + 1; stored, recorded
    const bool yin = y < Ny - 1, zin = z < Nz - 1;
    T* const sH0 = sm;            // [2][3][NT] input H totals, by plane parity
    T* const sEA = sm + 6 * NT;   // [2][3][NT] E totals after step n
    T* const sHA = sm + 12 * NT;  // [2][3][NT] H totals after step n
    T* const sEB = sm + 18 * NT;  // [2][3][NT] E totals after step n + 1
    const Fields<T>& A = I.F;
    const Fields<T>& B = I.G;
    const int lo = max(x0 - 2, 0), aHi = min(x1 + 1, Nx);  // step n's E on planes [lo, aHi)
    const int hLo = max(x0 - 1, 0);                        // step n's H on [hLo, aHi), step n + 1's E on [hLo, x1)
    const uint32_t p = valid ? ((uint32_t)y * (uint32_t)Nz + (uint32_t)z) : 0u;
    const uint32_t base = ((uint32_t)lo * (uint32_t)Ny * (uint32_t)Nz + p) * S;
    auto off = [&](int x) { return base + (uint32_t)(x - lo) * sx; };
    auto load_h = [&](int x) {
        Split6<T> h{};
        if (valid && x < Nx) {
            const uint32_t b = off(x);
            h.xy = ldg((const G*)A.hx_y, b); h.xz = ldg((const G*)A.hx_z, b); h.yz = ldg((const G*)A.hy_z, b);
            h.yx = ldg((const G*)A.hy_x, b); h.zx = ldg((const G*)A.hz_x, b); h.zy = ldg((const G*)A.hz_y, b);
        }
        return h;
    };
    auto load_e = [&](int x) {
        Split6<T> e{};
        if (mAe && x < aHi) {
            const uint32_t b = off(x);
            e.xy = ldg((const G*)A.ex_y, b); e.xz = ldg((const G*)A.ex_z, b); e.yz = ldg((const G*)A.ey_z, b);
            e.yx = ldg((const G*)A.ey_x, b); e.zx = ldg((const G*)A.ez_x, b); e.zy = ldg((const G*)A.ez_y, b);
        }
        return e;
    };
    auto store = [&](uint32_t b, const Split6<T>& e, bool h) {
        if (h) {
            stg((G*)B.hx_y, b, e.xy); stg((G*)B.hx_z, b, e.xz); stg((G*)B.hy_z, b, e.yz);
            stg((G*)B.hy_x, b, e.yx); stg((G*)B.hz_x, b, e.zx); stg((G*)B.hz_y, b, e.zy);
        } else {
            stg((G*)B.ex_y, b, e.xy); stg((G*)B.ex_z, b, e.xz); stg((G*)B.ey_z, b, e.yz);
            stg((G*)B.ey_x, b, e.yx); stg((G*)B.ez_x, b, e.zx); stg((G*)B.ez_y, b, e.zy);
        }
    };
    auto put3 = [&](T* buf, T a, T b, T c) {
        buf[tid] = a;
        buf[NT + tid] = b;
        buf[2 * NT + tid] = c;
    };
    const bool eprobes = io && I.negroups, hprobes = io && I.nhgroups;
    auto cell_of = [&](int x) { return ((uint32_t)x * (uint32_t)Ny + (uint32_t)y) * (uint32_t)Nz + (uint32_t)z; };
    auto ewd = [&](int x) { return eprobes && mAe && x < aHi ? ldg((const G32*)I.ebits, (cell_of(x) >> 5) * 4u) : 0u; };
    auto hwd = [&](int x) { return hprobes && mAh && x < aHi ? ldg((const G32*)I.hbits, (cell_of(x) >> 5) * 4u) : 0u; };
    auto bit = [&](uint32_t word, int x) { return ((word >> (cell_of(x) & 31u)) & 1u) != 0; };
    auto ekey = [&](int x) { return mAe && x < aHi ? C.ekey(off(x)) : 0u; };
    auto hkey = [&](int x) { return mAh && x < aHi ? C.hkey(off(x)) : 0u; };

    // input H splits of planes i - 1, i, i + 1 and E splits of plane i
    Split6<T> h0p{}, h0c = load_h(lo), hN = load_h(lo + 1), e0c = load_e(lo);
    // the step-n states of planes i - 1, i - 2 (E) and i - 2, i - 3 (H) that the later phases advance
    Split6<T> eA1{}, eA2{}, hA2{}, hA3{};
    T eB3x = T(0), eB3y = T(0), eB3z = T(0), eB4y = T(0), eB4z = T(0);  // E totals after step n + 1, planes i - 3, i - 4
    // coefficient keys and probe bitmask words by plane: E of planes i, i - 1, i - 2, H of i .. i - 3
    uint32_t ek0 = ekey(lo), ek1 = 0u, ek2 = 0u, hk0 = hkey(lo), hk1 = 0u, hk2 = 0u, hk3 = 0u;
    uint32_t ew0 = ewd(lo), ew1 = 0u, ew2 = 0u, hw0 = hwd(lo), hw1 = 0u, hw2 = 0u, hw3 = 0u;
    for (int i = lo; i < x1 + 3; ++i) {
        const int p2 = i - 1, p3 = i - 2, p4 = i - 3;
        const bool d1 = i < aHi;                   // step n, E, plane i
        const bool d2 = p2 >= hLo && p2 < aHi;     // step n, H, plane i - 1
        const bool d3 = p3 >= hLo && p3 < x1;      // step n + 1, E, plane i - 2
        const bool d4 = p4 >= x0 && p4 < x1;       // step n + 1, H, plane i - 3
        const int par = i & 1, prv = par ^ 1;
        ECoef<T> c1{}, c3{};
        HCoef<T> c2{}, c4{};
        if (d1 && mAe) c1 = C.e_at(ek0);
        if (d2 && mAh) c2 = C.h_at(hk1);
        if (d3 && mBe) c3 = C.e_at(ek2);
        if (d4 && own) c4 = C.h_at(hk3);
        const Split6<T> hNN = load_h(i + 2);
        const Split6<T> eN = load_e(i + 1);
        const uint32_t ekN = ekey(i + 1), hkN = hkey(i + 1), ewN = ewd(i + 1), hwN = hwd(i + 1);
        const T h0x = h0c.xy + h0c.xz, h0y = h0c.yz + h0c.yx, h0z = h0c.zx + h0c.zy;
        put3(sH0 + par * 3 * NT, h0x, h0y, h0z);
        __syncthreads();
        // ---- step n, E, plane i (input H totals of plane i exchanged just now)
        Split6<T> eA = e0c;
        T eAx = T(0), eAy = T(0), eAz = T(0);
        if (d1 && mAe) {
            const T* const s0 = sH0 + par * 3 * NT;
            const bool xin = i < Nx - 1;
            const T hz_yn = yin ? s0[2 * NT + tid + 64] : T(0), hx_yn = yin ? s0[tid + 64] : T(0);
            const T hy_zn = zin ? s0[NT + tid + 1] : T(0), hx_zn = zin ? s0[tid + 1] : T(0);
            const T hy_xn = xin ? hN.yz + hN.yx : T(0), hz_xn = xin ? hN.zx + hN.zy : T(0);
            if (yin) eA.xy = c1.ca_xy * eA.xy + c1.cb_xy * (hz_yn - h0z);
            if (zin) eA.xz = c1.ca_xz * eA.xz + c1.cb_xz * (hy_zn - h0y);
            if (zin) eA.yz = c1.ca_yz * eA.yz + c1.cb_yz * (hx_zn - h0x);
            if (xin) eA.yx = c1.ca_yx * eA.yx + c1.cb_yx * (hz_xn - h0z);
            if (xin) eA.zx = c1.ca_zx * eA.zx + c1.cb_zx * (hy_xn - h0y);
            if (yin) eA.zy = c1.ca_zy * eA.zy + c1.cb_zy * (hx_yn - h0x);
            const bool rec = own && i >= x0 && i < x1;
            if (eprobes && bit(ew0, i)) inject_e(I, cell_of(i), n, rec, eA.xy, eA.xz, eA.yz, eA.yx, eA.zx, eA.zy);
            eAx = eA.xy + eA.xz;
            eAy = eA.yz + eA.yx;
            eAz = eA.zx + eA.zy;
            if (rec && nmon) p2_sample(mons, mlist, nmon, EX, i, y, z, n, slot, eAx, eAy, eAz);
        }
        put3(sEA + par * 3 * NT, eAx, eAy, eAz);
        // ---- step n, H, plane i - 1 (its E totals were exchanged an iteration ago)
        Split6<T> hA = h0p;
        T hAx = T(0), hAy = T(0), hAz = T(0);
        if (d2 && mAh) {
            const T* const s1 = sEA + prv * 3 * NT;
            const T ex = eA1.xy + eA1.xz, ey = eA1.yz + eA1.yx, ez = eA1.zx + eA1.zy;
            const T ey_xm = p2 > 0 ? eA2.yz + eA2.yx : T(0), ez_xm = p2 > 0 ? eA2.zx + eA2.zy : T(0);
            const T ez_ym = y > 0 ? s1[2 * NT + tid - 64] : T(0), ex_ym = y > 0 ? s1[tid - 64] : T(0);
            const T ey_zm = z > 0 ? s1[NT + tid - 1] : T(0), ex_zm = z > 0 ? s1[tid - 1] : T(0);
            hA.xy = c2.da_xy * h0p.xy + (c2.db_xy2 * ez - c2.db_xy1 * ez_ym);
            hA.xz = c2.da_xz * h0p.xz + (c2.db_xz2 * ey - c2.db_xz1 * ey_zm);
            hA.yz = c2.da_yz * h0p.yz + (c2.db_yz2 * ex - c2.db_yz1 * ex_zm);
            hA.yx = c2.da_yx * h0p.yx + (c2.db_yx2 * ez - c2.db_yx1 * ez_xm);
            hA.zx = c2.da_zx * h0p.zx + (c2.db_zx2 * ey - c2.db_zx1 * ey_xm);
            hA.zy = c2.da_zy * h0p.zy + (c2.db_zy2 * ex - c2.db_zy1 * ex_ym);
            const bool rec = own && p2 >= x0 && p2 < x1;
            if (hprobes && bit(hw1, p2))
                inject_cell((const ProbeGroup<T>*)I.hgroups, I.nhgroups, cell_of(p2), n, rec, hA.xy, hA.xz, hA.yz,
                            hA.yx, hA.zx, hA.zy);
            hAx = hA.xy + hA.xz;
            hAy = hA.yz + hA.yx;
            hAz = hA.zx + hA.zy;
            if (rec && nmon) p2_sample(mons, mlist, nmon, HX, p2, y, z, n, slot, hAx, hAy, hAz);
        }
        put3(sHA + par * 3 * NT, hAx, hAy, hAz);
        // ---- step n + 1, E, plane i - 2 (H totals of plane i - 2 after step n: an iteration ago;
        // those of plane i - 1 at this cell: just computed)
        T eBx = T(0), eBy = T(0), eBz = T(0);
        if (d3 && mBe) {
            const T* const s2 = sHA + prv * 3 * NT;
            Split6<T> eB = eA2;
            const bool xin = p3 < Nx - 1;
            const T hx_c = hA2.xy + hA2.xz, hy_c = hA2.yz + hA2.yx, hz_c = hA2.zx + hA2.zy;
            const T hz_yn = yin ? s2[2 * NT + tid + 64] : T(0), hx_yn = yin ? s2[tid + 64] : T(0);
            const T hy_zn = zin ? s2[NT + tid + 1] : T(0), hx_zn = zin ? s2[tid + 1] : T(0);
            const T hy_xn = xin ? hAy : T(0), hz_xn = xin ? hAz : T(0);
            if (yin) eB.xy = c3.ca_xy * eB.xy + c3.cb_xy * (hz_yn - hz_c);
            if (zin) eB.xz = c3.ca_xz * eB.xz + c3.cb_xz * (hy_zn - hy_c);
            if (zin) eB.yz = c3.ca_yz * eB.yz + c3.cb_yz * (hx_zn - hx_c);
            if (xin) eB.yx = c3.ca_yx * eB.yx + c3.cb_yx * (hz_xn - hz_c);
            if (xin) eB.zx = c3.ca_zx * eB.zx + c3.cb_zx * (hy_xn - hy_c);
            if (yin) eB.zy = c3.ca_zy * eB.zy + c3.cb_zy * (hx_yn - hx_c);
            const bool rec = own && p3 >= x0;
            if (eprobes && bit(ew2, p3))
                inject_e(I, cell_of(p3), n + 1, rec, eB.xy, eB.xz, eB.yz, eB.yx, eB.zx, eB.zy);
            eBx = eB.xy + eB.xz;
            eBy = eB.yz + eB.yx;
            eBz = eB.zx + eB.zy;
            if (rec) store(off(p3), eB, false);
        }
        put3(sEB + par * 3 * NT, eBx, eBy, eBz);
        // ---- step n + 1, H, plane i - 3 (its E totals were exchanged an iteration ago)
        if (d4 && own) {
            const T* const s3 = sEB + prv * 3 * NT;
            const T ey_xm = p4 > 0 ? eB4y : T(0), ez_xm = p4 > 0 ? eB4z : T(0);
            const T ez_ym = y > 0 ? s3[2 * NT + tid - 64] : T(0), ex_ym = y > 0 ? s3[tid - 64] : T(0);
            const T ey_zm = z > 0 ? s3[NT + tid - 1] : T(0), ex_zm = z > 0 ? s3[tid - 1] : T(0);
            Split6<T> hB;
            hB.xy = c4.da_xy * hA3.xy + (c4.db_xy2 * eB3z - c4.db_xy1 * ez_ym);
            hB.xz = c4.da_xz * hA3.xz + (c4.db_xz2 * eB3y - c4.db_xz1 * ey_zm);
            hB.yz = c4.da_yz * hA3.yz + (c4.db_yz2 * eB3x - c4.db_yz1 * ex_zm);
            hB.yx = c4.da_yx * hA3.yx + (c4.db_yx2 * eB3z - c4.db_yx1 * ez_xm);
            hB.zx = c4.da_zx * hA3.zx + (c4.db_zx2 * eB3y - c4.db_zx1 * ey_xm);
            hB.zy = c4.da_zy * hA3.zy + (c4.db_zy2 * eB3x - c4.db_zy1 * ex_ym);
            if (hprobes && bit(hw3, p4))
                inject_cell((const ProbeGroup<T>*)I.hgroups, I.nhgroups, cell_of(p4), n + 1, true, hB.xy, hB.xz,
                            hB.yz, hB.yx, hB.zx, hB.zy);
            store(off(p4), hB, true);
        }
        eB4y = eB3y;
        eB4z = eB3z;
        eB3x = eBx;
        eB3y = eBy;
        eB3z = eBz;
        hA3 = hA2;
        hA2 = hA;
        eA2 = eA1;
        eA1 = eA;
        h0p = h0c;
        h0c = hN;
        hN = hNN;
        e0c = eN;
        ek2 = ek1; ek1 = ek0; ek0 = ekN;
        hk3 = hk2; hk2 = hk1; hk1 = hk0; hk0 = hkN;
        ew2 = ew1; ew1 = ew0; ew0 = ewN;
        hw3 = hw2; hw2 = hw1; hw1 = hw0; hw0 = hwN;
    }
}

// two != 0: steps n and n + 1 (n from base/local); the phasor monitors' sample rows of step n's
// staging slot are written by logical workgroup 0 (k_monitors writes step n + 1's).
template <typename T, int TY, bool SKEW, int RS = 64>
__global__ void __launch_bounds__((TY + 4) * RS) k_step2(const InstDev<T>* __restrict__ insts,
                                                            const MemberDev* __restrict__ mem, int nmem, int total,
                                                            const int* base, int local, int io, int two,
                                                            const MonitorDev* __restrict__ mons, int nmon, int slot,
                                                            int* __restrict__ rows) {
    __shared__ int s_member;
    __shared__ T sm[(SKEW ? 24 : 15) * (TY + 4) * RS];
    constexpr int FZ = RS - 4;  // z cells owned per row
    const int L = xcd_logical(blockIdx.x, gridDim.x);
    if (threadIdx.x == 0) {
        int m = 0;
        if (L < total)
            while (m + 1 < nmem && mem[m + 1].tile0 <= L) ++m;
        s_member = m;
    }
    __syncthreads();
    if (L >= total) return;  // uniform over the workgroup: no barrier is skipped by part of it
    const int n = step_of(base, local);
    const int nm = (two && io) ? nmon : 0;
    if (nm && L == 0 && (int)threadIdx.x < nm) {
        const MonitorDev& M = mons[threadIdx.x];
        if (M.nf) rows[M.id * kStageSlots + slot] = n % M.n_step == 0 ? n / M.n_step : -1;
    }
    const int m = __builtin_amdgcn_readfirstlane(s_member);
    const MemberDev M = mem[m];
    const int local_t = L - M.tile0;
    const int e = local_t % M.K, tile = local_t / M.K;
    const int inst = __builtin_amdgcn_readfirstlane(M.inst0 + e);
    const InstDev<T>* I = insts + inst;
    const int tz_n = (I->Nz + FZ - 1) / FZ;
    const int pb = tile % I->npb, xc = tile / I->npb;
    const int y0 = (pb / tz_n) * TY, z0 = (pb % tz_n) * FZ;
    const int x0 = xc * I->xplanes, x1 = min(I->Nx, x0 + I->xplanes);
    // the monitors this tile samples at step n: of this instance, sampling at n, with their plane
    // through the tile's owned cells (usually none); wavefront 0 compacts them into LDS
    __shared__ int s_mlist[kP2Mon];
    __shared__ int s_mcount;
    int nmt = 0;
    if (nm) {  // uniform over the workgroup
        if (threadIdx.x < 64) {
            const int k = (int)threadIdx.x;
            bool hit = false;
            if (k < nm) {
                const MonitorDev& Mk = mons[k];
                if (Mk.inst == inst && n % Mk.n_step == 0)
                    hit = Mk.axis == 0 ? (Mk.pos >= x0 && Mk.pos < x1)
                        : Mk.axis == 1 ? (Mk.pos >= y0 && Mk.pos < y0 + TY)
                                       : (Mk.pos >= z0 && Mk.pos < z0 + FZ);
            }
            const uint64_t ball = __ballot(hit);
            if (hit) s_mlist[__popcll(ball & ((1ull << k) - 1ull))] = k;
            if (k == 0) s_mcount = __popcll(ball);
        }
        __syncthreads();
        nmt = __builtin_amdgcn_readfirstlane(s_mcount);
    }
    static_assert(!SKEW || RS == 64, "the skewed pass has 64-lane rows");
    if (SKEW && two) {
        if (I->coef_kind == 1)
            two_step_skew_tile<T, TY, CoefTable<T>>(*I, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, x0, x1, y0,
                                                    z0, n, io != 0, mons, s_mlist, nmt, slot, sm);
        else
            two_step_skew_tile<T, TY, CoefRaw<T>>(*I, CoefRaw<T>{I->raw}, x0, x1, y0, z0, n, io != 0, mons, s_mlist, nmt,
                                                  slot, sm);
    } else if (I->coef_kind == 1) {
        two_step_tile<T, TY, CoefTable<T>, RS>(*I, CoefTable<T>{I->eidx, I->hidx, I->etab, I->htab}, x0, x1, y0, z0,
                                               n, io != 0, two != 0, mons, s_mlist, nmt, slot, sm);
    } else {
        two_step_tile<T, TY, CoefRaw<T>, RS>(*I, CoefRaw<T>{I->raw}, x0, x1, y0, z0, n, io != 0, two != 0, mons,
                                             s_mlist, nmt, slot, sm);
    }
}

// ------------------------------------------------------------------------------------------------
// far field (NF2FF): radiation integrals of the equivalent surface currents of a closed box
// ------------------------------------------------------------------------------------------------

struct FFFace {
    int32_t axis, n0, n1, pad;
    float pos;               // face position along its normal, metres
    int32_t cell0;           // first cell of this face in the concatenated cell list
    const float* r0;         // [n0][n1] cell centres along the first in-face axis
    const float* r1;         // [n0][n1] along the second
    const float* ds;         // [n0][n1] cell areas
    const float2* J;         // [3][nf][n0][n1] n x H
    const float2* M;         // [3][nf][n0][n1] -n x E
};

constexpr int kFFChunk = 256;

// One thread per direction, a workgroup per (frequency, block of directions); the face cells pass
// through LDS in chunks.  For every cell: I = exp(j beta r'.r^) dS, and
//   N_theta += (Jx ct cp + Jy ct sp - Jz st) I,  N_phi += (-Jx sp + Jy cp) I,  L likewise with M;
// then E_theta = (-j beta / 4 pi)(L_phi + eta0 N_theta), E_phi = (j beta / 4 pi)(L_theta - eta0 N_phi).
__global__ void __launch_bounds__(kFFChunk) k_nf2ff(const FFFace* __restrict__ faces, int nfaces, int ncells, int nf,
                                                   const float* __restrict__ beta, const float* __restrict__ theta,
                                                   int ntheta, const float* __restrict__ phi, int nphi, float eta0,
                                                   float2* __restrict__ out) {
    __shared__ float s_r0[kFFChunk], s_r1[kFFChunk], s_ds[kFFChunk];
    __shared__ float2 s_j[3][kFFChunk], s_m[3][kFFChunk];
    __shared__ int s_axis[kFFChunk];
    __shared__ float s_pos[kFFChunk];
    const int ndir = ntheta * nphi;
    const int dblocks = (ndir + kFFChunk - 1) / kFFChunk;
    const int f = blockIdx.x / dblocks;
    const int dir = (blockIdx.x % dblocks) * kFFChunk + threadIdx.x;
    const bool live = dir < ndir;
    const int th = live ? dir / nphi : 0, ph = live ? dir % nphi : 0;
    float st, ct, sp, cp;
    sincosf(theta[th], &st, &ct);
    sincosf(phi[ph], &sp, &cp);
    const float u = st * cp, v = st * sp, w = ct;
    const float b = beta[f];
    float2 Nt = make_float2(0.f, 0.f), Np = Nt, Lt = Nt, Lp = Nt;
    for (int c0 = 0; c0 < ncells; c0 += kFFChunk) {
        __syncthreads();
        const int c = c0 + threadIdx.x;
        if (c < ncells) {
            int k = 0;
            while (k + 1 < nfaces && faces[k + 1].cell0 <= c) ++k;
            const FFFace& F = faces[k];
            const int i = c - F.cell0;
            const int64_t plane = (int64_t)F.n0 * F.n1;
            s_r0[threadIdx.x] = F.r0[i];
            s_r1[threadIdx.x] = F.r1[i];
            s_ds[threadIdx.x] = F.ds[i];
            s_axis[threadIdx.x] = F.axis;
            s_pos[threadIdx.x] = F.pos;
            for (int a = 0; a < 3; ++a) {
                s_j[a][threadIdx.x] = F.J[((int64_t)a * nf + f) * plane + i];
                s_m[a][threadIdx.x] = F.M[((int64_t)a * nf + f) * plane + i];
            }
        }
        __syncthreads();
        const int n = min(kFFChunk, ncells - c0);
        for (int q = 0; q < n; ++q) {
            const int ax = s_axis[q];
            const float r0 = s_r0[q], r1 = s_r1[q];
            const float rdot = ax == 0 ? s_pos[q] * u + (r0 * v + r1 * w)
                             : ax == 1 ? s_pos[q] * v + (r0 * u + r1 * w)
                                       : s_pos[q] * w + (r0 * u + r1 * v);
            float si, co;
            sincosf(b * rdot, &si, &co);
            const float ire = co * s_ds[q], iim = si * s_ds[q];
            const float2 jx = s_j[0][q], jy = s_j[1][q], jz = s_j[2][q];
            const float2 mx = s_m[0][q], my = s_m[1][q], mz = s_m[2][q];
            // theta and phi projections of the currents (real factors of the direction)
            const float2 jt = make_float2(jx.x * (ct * cp) + jy.x * (ct * sp) - jz.x * st,
                                          jx.y * (ct * cp) + jy.y * (ct * sp) - jz.y * st);
            const float2 jp = make_float2(-jx.x * sp + jy.x * cp, -jx.y * sp + jy.y * cp);
            const float2 mt = make_float2(mx.x * (ct * cp) + my.x * (ct * sp) - mz.x * st,
                                          mx.y * (ct * cp) + my.y * (ct * sp) - mz.y * st);
            const float2 mp = make_float2(-mx.x * sp + my.x * cp, -mx.y * sp + my.y * cp);
            Nt.x += jt.x * ire - jt.y * iim; Nt.y += jt.x * iim + jt.y * ire;
            Np.x += jp.x * ire - jp.y * iim; Np.y += jp.x * iim + jp.y * ire;
            Lt.x += mt.x * ire - mt.y * iim; Lt.y += mt.x * iim + mt.y * ire;
            Lp.x += mp.x * ire - mp.y * iim; Lp.y += mp.x * iim + mp.y * ire;
        }
    }
    if (!live) return;
    const float k4 = b / (4.0f * 3.14159265358979f);
    // -j k4 (Lp + eta0 Nt) and j k4 (Lt - eta0 Np)
    const float2 a1 = make_float2(Lp.x + eta0 * Nt.x, Lp.y + eta0 * Nt.y);
    const float2 a2 = make_float2(Lt.x - eta0 * Np.x, Lt.y - eta0 * Np.y);
    out[((int64_t)0 * nf + f) * ndir + dir] = make_float2(k4 * a1.y, -k4 * a1.x);
    out[((int64_t)1 * nf + f) * ndir + dir] = make_float2(-k4 * a2.y, k4 * a2.x);
}

__global__ void k_set_step(int* p, int v) { *p = v; }

// Plain streaming copy (16 B per lane per iteration): the measured bandwidth ceiling, and the
// reference point that scales the profiler's HBM byte counters.
__global__ void __launch_bounds__(256) k_stream_copy(float4* __restrict__ dst, const float4* __restrict__ src,
                                                     int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Device-wide barrier probe: what one time step would pay to stay inside a single persistent launch
// instead of ending at a kernel boundary.  The workgroups of a cooperative launch meet `iters` times
// at a flag-gather barrier: each stores its epoch in a 128-byte slot of its own, workgroup 0 waits for
// every slot and publishes the epoch, the others wait for that word.  FENCES 0: no fences (the
// synchronisation alone; stores made before it are then not guaranteed visible across the XCDs'
// L2 caches); 1: one wave per workgroup makes its XCD's L2 write back before arriving and
// invalidates it after leaving (agent-scope release and acquire); 2: every wave does.  Each wait is
// bounded by `limit` wall-clock ticks; one that runs out sets the abort word, and every wait also
// ends when the abort word is set, so all workgroups leave.
// words: [0] published epoch, [32] abort, [64 + 32 g] workgroup g's epoch.
template <int FENCES>
__global__ void __launch_bounds__(256) k_barrier_probe(uint32_t* words, int iters, long long limit) {
    __shared__ int s_abort;
    uint32_t* const pub = words;
    uint32_t* const abort_word = words + 32;
    const uint32_t G = gridDim.x;
    auto wait_ge = [&](const uint32_t* w, uint32_t target) {
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
            if (wall_clock64() - t0 > limit) {
                __hip_atomic_store(abort_word, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        return true;
    };
    if (threadIdx.x == 0) s_abort = 0;
    for (int k = 1; k <= iters; ++k) {
        const uint32_t epoch = (uint32_t)k;
        if (FENCES == 2) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (FENCES == 1 && threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (threadIdx.x == 0)
            __hip_atomic_store(words + 64 + 32 * blockIdx.x, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool ok = true;
        if (blockIdx.x == 0) {
            for (uint32_t g = threadIdx.x; g < G && ok; g += blockDim.x) ok = wait_ge(words + 64 + 32 * g, epoch);
            if (!ok) s_abort = 1;
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(pub, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (threadIdx.x == 0) {
            if (!wait_ge(pub, epoch)) s_abort = 1;
        }
        if (FENCES == 1 && threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
        if (FENCES == 2) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (s_abort) return;  // uniform over the workgroup (read after the barrier)
    }
}

// An empty kernel: graph-replayed launches of it give the spacing of dependent kernel boundaries.
__global__ void __launch_bounds__(256) k_boundary_probe() {}

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// A grid and its coefficients, shared by the member's instances.
struct Member {
    int Nx = 0, Ny = 0, Nz = 0;
    int64_t cells = 0;
    int K = 1, inst0 = 0;
    int xplanes_req = 0, xplanes = 0;  // requested (0 = automatic) and effective planes per workgroup
    void* d_coef = nullptr;
    size_t coef_bytes = 0;
    int coef_kind = 0;  // 0 raw, 1 table
    int64_t nE = 0, nH = 0;
    const void* raw[30] = {};
    const uint32_t* eidx = nullptr;
    const uint32_t* hidx = nullptr;
    const void* etab = nullptr;
    const void* htab = nullptr;
    // design-region coefficients assembled on the device (table form): the E table grows to hold
    // one entry per cell of each design box after the nE deduplicated ones; updates of the same box
    // rewrite its entries in place
    void* d_etab_design = nullptr;
    int64_t etab_cap = 0, design_used = 0, design_base = -1;
    int32_t design_box[6] = {};
    // a design box prepared for updates from device memory (rfdtd_design_prepare): its cell
    // permittivity (updated in place), conductivity, widths, inverses and coefficient pointers, held
    // in one device allocation
    struct DesignHeld {
        DesignDev D{};
        double* eps = nullptr;
        void* const* raw = nullptr;
        void* buf = nullptr;
        int64_t base = 0;
        bool ready = false;
    } held;
};

struct MonitorHost {
    rfdtd_monitor desc;
    bool active = false;
    int N1 = 0, N2 = 0, Nm = 0, pos = 0;
    int ring = 0;  // streamed time-domain monitor: sample planes kept on the device (0: all Nm)
    void* d_values = nullptr;
    void* d_phase = nullptr;
};

// One excitation of one member: its fields, probe series and monitors.
struct Instance {
    int member = 0;
    void* d_fields = nullptr;
    void* d_fields2 = nullptr;  // the second buffer of the fused step's pair
    size_t fields_bytes = 0, pitch = 0;
    uint32_t* d_ebits = nullptr;  // E-probe cells (one bit per cell)
    void* d_egroups = nullptr;    // E probe groups sorted by cell
    int negroups = 0;
    uint32_t* d_hbits = nullptr;  // the same for the H probes
    void* d_hgroups = nullptr;
    int nhgroups = 0;
    // probes
    int P = 0;
    void* d_series = nullptr;
    void* d_series_in = nullptr;  // pristine source samples, for recomputation
    int32_t* d_order = nullptr;
    int32_t* d_is_src = nullptr;
    std::vector<void*> probe_host_values;
    struct HostGroup {
        int64_t cell;
        int field, start, count;
    };
    std::vector<HostGroup> groups;  // E groups first, then H
    int ngE = 0;
    // monitors
    std::vector<MonitorHost> mons;
    // checkpoints: the 12 split arrays at steps 0, every, 2*every, ...
    std::vector<void*> ckpt;
};

struct Session {
    bool f64 = false;
    size_t esz = 4;
    uint32_t flags = 0;
    hipStream_t own_stream = nullptr;
    std::vector<Member> members;
    std::vector<Instance> insts;
    int Nt_probes = 0, Nt_monitors = 0;

    // device tables rebuilt by refresh()
    void* d_insts = nullptr;
    MemberDev* d_members = nullptr;
    void* d_groups = nullptr;  // E groups of all instances, then H groups
    int ngE = 0, ngH = 0;
    MonitorDev* d_mons = nullptr;
    int n_active = 0, mon_blocks = 0, flush_blocks = 0;
    void* d_stage = nullptr;  // staged phasor-monitor samples
    int* d_rows = nullptr;    // [n_active][kStageSlots] phase row of each staged sample
    int total_tiles = 0;
    bool dirty = true;

    int64_t step = 0;
    int block = 256, graph_steps = 100;
    bool pipelined = true;  // issue the next x plane's loads before this plane's stores
    int ckpt_every = 0;     // save the split fields every this many steps (0: off)
    bool tuning = false;    // autotune trials: field updates only, no probes or monitors
    int fused = 0;          // 0: separate E and H kernels; else the fused step with this many rows per tile
    bool nontemporal = false;  // separate pipelined kernels with nontemporal field stores
    bool nt_loads = false;     // ... and nontemporal loads of the arrays streamed once (variant 10)
    int vec = 1;               // separate kernels over vec consecutive z cells per thread (variants 11, 12)
    int p2_rs = 64;            // two-step tiles: lanes per row (64, or 32 for variants 13, 14)
    // the separate kernels' own launch shape (threads, x planes per workgroup, 0 = automatic), which
    // a gradient's recomputation uses whatever the forward variant: the tuner's best of variants
    // 0, 1, 4, 10
    int sep_block = 256, sep_xplanes = 0;
    // workgroup numbering of the column kernels (MemberDev::order), forward steps and recomputation
    int col_order = 0, sep_order = 0;
    bool fused_deep = false;   // fused step with E one plane ahead of H (one barrier per plane)
    bool two_step = false;     // two-step pass (k_step2): pairs of steps per launch; fused = its rows
    bool two_skew = false;     // the two-step pass's one-barrier form (variant 9)
    // streamed time-domain monitors: a device ring of stream_ring sample planes per monitor, copied
    // to the caller's host arrays on copy_stream after every run segment
    int stream_ring = 0;
    hipStream_t copy_stream = nullptr;
    hipEvent_t seg_ev = nullptr, cpy_ev[2] = {nullptr, nullptr};
    bool cpy_pending[2] = {false, false};
    int64_t seg_count = 0;
    // reverse-pass launch shape: pipelined (1) or plain (0) columns, threads per workgroup and x
    // chunks per grid (0: the forward's planes per workgroup); adj_block 0 = timed at the first
    // gradient and kept
    int adj_pf = 1, adj_block = 0, adj_chunks = 0;
    int parity = 0;         // which buffer of each instance's pair holds the current fields
    bool totals_stale = false;  // the fused step leaves the total arrays unwritten
    void* d_insts2 = nullptr;   // descriptors with the buffers of each pair exchanged
    // design-box H history: recorded every step of the forward run so that a gradient needs no
    // recomputation (per item: instance, box, buffer of capacity steps x slot_elems values)
    struct History {
        int inst;
        int32_t box[6];
        void* d;
        int64_t slot_elems, capacity;
    };
    std::vector<History> hist;
    void* d_hist_boxes = nullptr;
    int64_t* d_hist_h0 = nullptr;
    int hist_blocks = 0;
    int64_t* d_hist_hb = nullptr;  // the save's workgroup offsets in workgroups of the E launch's size
    int hist_blocks_e = 0;
    std::vector<int64_t> flat0; // block offsets of a flat launch over all instances' cells
    int64_t* d_flat0 = nullptr;

    int* d_step_base = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // captured chunk, per starting parity
    hipGraph_t graph_def[2] = {nullptr, nullptr};
    bool graph_dirty = true;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // ev1: the end of the last run, on the stream it ran on
    // a design update from device memory on some stream: the next run, on any stream, waits for it
    hipEvent_t design_ev = nullptr;
    bool design_pending = false;
    double last_total_ms = 0, last_e_ms = 0, last_h_ms = 0;
    static constexpr int kTimedSteps = 64;
    std::vector<hipEvent_t> kev;  // 4 per timed step: E start/end, H start/end
    int n_timed = 0;

    ~Session();
    void drop_graph();
    void release_probes(Instance& I);
    void release_monitors(Instance& I);
    // the buffer that holds the instance's current fields
    void* cur(const Instance& I) const { return parity ? I.d_fields2 : I.d_fields; }
};

void Session::drop_graph() {
    // replays of these graphs may still be queued (a launch-shape or design change between runs
    // does not synchronize): drain the device before their executables go
    if (graph[0] || graph[1]) (void)hipDeviceSynchronize();
    for (int p = 0; p < 2; ++p) {
        if (graph[p]) (void)hipGraphExecDestroy(graph[p]);
        if (graph_def[p]) (void)hipGraphDestroy(graph_def[p]);
        graph[p] = nullptr;
        graph_def[p] = nullptr;
    }
    graph_dirty = true;
}

void Session::release_probes(Instance& I) {
    if (I.d_series) (void)hipFree(I.d_series);
    if (I.d_series_in) (void)hipFree(I.d_series_in);
    I.d_series_in = nullptr;
    if (I.d_order) (void)hipFree(I.d_order);
    if (I.d_is_src) (void)hipFree(I.d_is_src);
    I.d_series = nullptr; I.d_order = nullptr; I.d_is_src = nullptr;
    for (void* p : {(void*)I.d_ebits, I.d_egroups, (void*)I.d_hbits, I.d_hgroups})
        if (p) (void)hipFree(p);
    I.d_ebits = nullptr; I.d_egroups = nullptr; I.negroups = 0;
    I.d_hbits = nullptr; I.d_hgroups = nullptr; I.nhgroups = 0;
    I.P = 0; I.ngE = 0;
    I.groups.clear();
    I.probe_host_values.clear();
}

void Session::release_monitors(Instance& I) {
    for (auto& m : I.mons) {
        if (m.d_values) (void)hipFree(m.d_values);
        if (m.d_phase) (void)hipFree(m.d_phase);
    }
    I.mons.clear();
}

Session::~Session() {
    (void)hipDeviceSynchronize();
    drop_graph();
    for (auto& I : insts) {
        release_probes(I);
        release_monitors(I);
        if (I.d_fields) (void)hipFree(I.d_fields);
        if (I.d_fields2) (void)hipFree(I.d_fields2);
        for (void* c : I.ckpt)
            if (c) (void)hipFree(c);
    }
    for (auto& M : members) {
        if (M.d_coef) (void)hipFree(M.d_coef);
        if (M.d_etab_design) (void)hipFree(M.d_etab_design);
        if (M.held.buf) (void)hipFree(M.held.buf);
    }
    for (auto& h : hist)
        if (h.d) (void)hipFree(h.d);
    for (void* p : {(void*)d_insts, (void*)d_members, d_groups, (void*)d_mons, (void*)d_step_base, d_stage, (void*)d_rows,
                    d_insts2, (void*)d_flat0, d_hist_boxes, (void*)d_hist_h0, (void*)d_hist_hb})
        if (p) (void)hipFree(p);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    for (hipEvent_t e : {seg_ev, cpy_ev[0], cpy_ev[1]})
        if (e) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (design_ev) (void)hipEventDestroy(design_ev);
    for (auto e : kev) (void)hipEventDestroy(e);
    if (own_stream) (void)hipStreamDestroy(own_stream);
}

// ---- bit-exact deduplication of coefficient tuples ----

__host__ __device__ inline uint64_t mix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}

template <typename U, int K>
__host__ __device__ inline uint64_t tuple_hash(const U* k) {
    uint64_t h = 0x9e3779b97f4a7c15ULL;
    for (int j = 0; j < K; ++j) h = mix64(h ^ (uint64_t)k[j]) + (uint64_t)j;
    return h;
}

// Open-addressing set of distinct K-tuples (compared bit for bit), numbered in insertion order.
template <typename U, int K>
struct TupleSet {
    std::vector<U> table;  // K values per entry
    std::vector<int32_t> slots = std::vector<int32_t>(1 << 12, -1);
    int64_t count = 0;
    static uint64_t hash(const U* k) { return tuple_hash<U, K>(k); }
    // entry of key, inserted if new; -1 once more than max_entries would exist
    int32_t find_or_insert(const U* key, int64_t max_entries) {
        size_t cap = slots.size();
        size_t s = hash(key) & (cap - 1);
        while (slots[s] >= 0) {
            const U* e = &table[(size_t)slots[s] * K];
            bool eq = true;
            for (int k = 0; k < K; ++k)
                if (e[k] != key[k]) { eq = false; break; }
            if (eq) return slots[s];
            s = (s + 1) & (cap - 1);
        }
        if (count >= max_entries) return -1;
        const int32_t id = (int32_t)count++;
        table.insert(table.end(), key, key + K);
        slots[s] = id;
        if ((size_t)count * 2 > cap) {  // grow and rehash
            cap <<= 1;
            std::vector<int32_t> ns(cap, -1);
            for (int64_t e = 0; e < count; ++e) {
                size_t t = hash(&table[(size_t)e * K]) & (cap - 1);
                while (ns[t] >= 0) t = (t + 1) & (cap - 1);
                ns[t] = (int32_t)e;
            }
            slots.swap(ns);
        }
        return id;
    }
};

// Collect the distinct K-tuples of arrays[0..K) over all cells into idx (per cell) and table.
// zero_last_x marks arrays whose x = Nx-1 plane reads as zero.  Gives up (returns false) once more
// than max_entries distinct tuples exist.  The cells are split into contiguous ranges deduplicated
// by separate host threads; merging their sets in range order numbers every tuple by its first
// cell, exactly as one pass over all cells would.
template <typename U, int K>
bool dedup(const U* const* arrays, const bool* zero_last_x, int64_t cells, int64_t plane, int64_t max_entries,
           std::vector<uint32_t>& idx, std::vector<U>& table) {
    idx.resize(cells);
    table.clear();
    const int64_t last0 = cells - plane;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)hw, 16, cells / (1 << 16)}));
    std::vector<TupleSet<U, K>> local(nt);
    std::vector<char> ok(nt, 1);
    auto range = [&](int t) {
        return std::make_pair(cells * t / nt, cells * (t + 1) / nt);
    };
    auto work = [&](int t) {
        const auto r = range(t);
        U key[K];
        for (int64_t i = r.first; i < r.second; ++i) {
            const bool last = i >= last0;
            for (int k = 0; k < K; ++k) key[k] = (last && zero_last_x[k]) ? U(0) : arrays[k][i];
            const int32_t id = local[t].find_or_insert(key, max_entries);
            if (id < 0) {
                ok[t] = 0;
                return;
            }
            idx[i] = (uint32_t)id;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (int t = 0; t < nt; ++t)
        if (!ok[t]) return false;
    // merge in range order; remap each range's local numbers to the global ones
    TupleSet<U, K>& all = local[0];
    std::vector<std::vector<uint32_t>> remap(nt);
    for (int t = 1; t < nt; ++t) {
        remap[t].resize(local[t].count);
        for (int64_t e = 0; e < local[t].count; ++e) {
            const int32_t id = all.find_or_insert(&local[t].table[(size_t)e * K], max_entries);
            if (id < 0) return false;
            remap[t][e] = (uint32_t)id;
        }
    }
    pool.clear();
    for (int t = 1; t < nt; ++t)
        pool.emplace_back([&, t] {
            const auto r = range(t);
            for (int64_t i = r.first; i < r.second; ++i) idx[i] = remap[t][idx[i]];
        });
    for (auto& th : pool) th.join();
    table.swap(all.table);
    return true;
}

// which of the 30 arrays read as zero on the x = Nx-1 plane (solver.cpp:294-328: Cy, Cz, Dx)
bool zero_last_plane(int k) { return (k >= 4 && k <= 11) || (k >= 12 && k <= 17); }

// ---- the same deduplication on the device ----
//
// Every cell's tuple is hashed; the (hash, cell) pairs are radix-sorted, so equal tuples sit together
// with their cells ascending; a run of equal tuples (compared bit for bit, so a hash collision only
// splits a run) is one group whose first cell is its first occurrence.  Sorting the groups by first
// cell and numbering them in that order gives each tuple the number the host pass gives it (first
// occurrence in cell order), which keeps neighbouring cells' entries close together in the table.

template <typename U, int K>
struct TuplePtrs {
    const U* p[K];
};

template <typename U, int K>
__device__ __forceinline__ void tuple_at(const TuplePtrs<U, K>& a, uint32_t c, U t[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) t[k] = a.p[k][c];
}

template <typename U, int K>
__global__ void __launch_bounds__(256) k_tuple_hash(TuplePtrs<U, K> a, uint32_t n, uint64_t* keys, uint32_t* cells) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    U t[K];
    tuple_at<U, K>(a, i, t);
    keys[i] = tuple_hash<U, K>(t);
    cells[i] = i;
}

// head[i] = 1 where sorted position i starts a run of equal tuples
template <typename U, int K>
__global__ void __launch_bounds__(256) k_tuple_heads(TuplePtrs<U, K> a, uint32_t n, const uint64_t* keys,
                                                     const uint32_t* cells, uint32_t* head) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h = 1;
    if (i > 0 && keys[i] == keys[i - 1]) {
        U t0[K], t1[K];
        tuple_at<U, K>(a, cells[i], t0);
        tuple_at<U, K>(a, cells[i - 1], t1);
        bool eq = true;
#pragma unroll
        for (int k = 0; k < K; ++k) eq &= t0[k] == t1[k];
        h = eq ? 0 : 1;
    }
    head[i] = h;
}

// first[g] = first cell of group g (grp: inclusive scan of the heads, 1-based)
__global__ void __launch_bounds__(256) k_group_first(uint32_t n, const uint32_t* head, const uint32_t* grp,
                                                     const uint32_t* cells, uint32_t* first, uint32_t* gid) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    first[grp[i] - 1] = cells[i];
    gid[grp[i] - 1] = grp[i] - 1;
}

// groups sorted by first cell: rank r is the entry number of group gid_sorted[r]; its tuple is the
// table entry (stride values per entry, the rest zero)
template <typename U, int K>
__global__ void __launch_bounds__(256) k_group_entries(TuplePtrs<U, K> a, uint32_t ng, const uint32_t* first_sorted,
                                                       const uint32_t* gid_sorted, uint32_t* entry_of, U* table,
                                                       int stride) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ng) return;
    entry_of[gid_sorted[r]] = r;
    U t[K];
    tuple_at<U, K>(a, first_sorted[r], t);
    for (int k = 0; k < stride; ++k) table[(size_t)r * stride + k] = k < K ? t[k] : U(0);
}

__global__ void __launch_bounds__(256) k_cell_entries(uint32_t n, const uint32_t* grp, const uint32_t* cells,
                                                      const uint32_t* entry_of, uint32_t* idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    idx[cells[i]] = entry_of[grp[i] - 1];
}

// Scratch of one device deduplication, freed on every exit path.
struct DedupScratch {
    std::vector<void*> ptrs;
    void* alloc(size_t bytes) {
        void* p = nullptr;
        HIP_OK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
        ptrs.push_back(p);
        return p;
    }
    ~DedupScratch() {
        (void)hipDeviceSynchronize();
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// Distinct K-tuples of the device arrays a over n cells, numbered by first occurrence.  Returns the
// number of entries; the per-cell entry numbers and the table (stride values per entry) are left in
// *idx and *table (device allocations owned by the caller) unless there are more than max_entries,
// in which case nothing is allocated.
template <typename U, int K>
int64_t device_dedup(const TuplePtrs<U, K>& a, uint32_t n, int64_t max_entries, int stride, uint32_t** idx,
                     U** table) {
    DedupScratch S;
    const int blocks = (int)((n + 255) / 256);
    uint64_t* keys = (uint64_t*)S.alloc((size_t)n * 8);
    uint64_t* keys_s = (uint64_t*)S.alloc((size_t)n * 8);
    uint32_t* cells = (uint32_t*)S.alloc((size_t)n * 4);
    uint32_t* cells_s = (uint32_t*)S.alloc((size_t)n * 4);
    uint32_t* head = (uint32_t*)S.alloc((size_t)n * 4);
    uint32_t* grp = (uint32_t*)S.alloc((size_t)n * 4);
    k_tuple_hash<U, K><<<blocks, 256>>>(a, n, keys, cells);
    HIP_OK(hipGetLastError());
    size_t tb = 0;
    HIP_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, keys_s, cells, cells_s, (int)n));
    void* tmp = S.alloc(tb);
    HIP_OK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, keys_s, cells, cells_s, (int)n));
    k_tuple_heads<U, K><<<blocks, 256>>>(a, n, keys_s, cells_s, head);
    HIP_OK(hipGetLastError());
    size_t tb2 = 0;
    HIP_OK(hipcub::DeviceScan::InclusiveSum(nullptr, tb2, head, grp, (int)n));
    void* tmp2 = S.alloc(tb2);
    HIP_OK(hipcub::DeviceScan::InclusiveSum(tmp2, tb2, head, grp, (int)n));
    uint32_t ng = 0;
    HIP_OK(hipMemcpy(&ng, grp + (n - 1), 4, hipMemcpyDeviceToHost));
    if ((int64_t)ng > max_entries) return ng;
    // reuse: first / gid per group, then their sorted copies
    uint32_t* first = (uint32_t*)keys;
    uint32_t* gid = first + ng;
    uint32_t* first_s = (uint32_t*)keys_s;
    uint32_t* gid_s = first_s + ng;
    k_group_first<<<blocks, 256>>>(n, head, grp, cells_s, first, gid);
    HIP_OK(hipGetLastError());
    size_t tb3 = 0;
    HIP_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, first, first_s, gid, gid_s, (int)ng));
    void* tmp3 = tb3 <= tb ? tmp : S.alloc(tb3);
    HIP_OK(hipcub::DeviceRadixSort::SortPairs(tmp3, tb3, first, first_s, gid, gid_s, (int)ng));
    uint32_t* entry_of = cells;  // ng <= n
    U* tab = nullptr;
    uint32_t* ix = nullptr;
    HIP_OK(hipMalloc(&tab, std::max<size_t>((size_t)ng * stride * sizeof(U), 256)));
    if (hipMalloc(&ix, (size_t)n * 4) != hipSuccess) {
        (void)hipFree(tab);
        RF_THROW(RFDTD_ENOMEM, "no device memory for the coefficient table");
    }
    k_group_entries<U, K><<<(int)((ng + 255) / 256), 256>>>(a, ng, first_s, gid_s, entry_of, tab, stride);
    k_cell_entries<<<blocks, 256>>>(n, grp, cells_s, entry_of, ix);
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(tab);
        (void)hipFree(ix);
        HIP_OK(e);
    }
    *idx = ix;
    *table = tab;
    return ng;
}

// Replace the member's raw device arrays by the table built from them on the device (numbered as
// the host pass numbers it), when it stays within the entry limit; else the raw arrays stay.
template <typename T>
void table_from_raw(Member& M, uint32_t flags) {
    using U = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    const int64_t cells = M.cells;
    const int64_t limit = (flags & RFDTD_COEF_TABLE) ? cells : std::max<int64_t>(4096, cells / 8);
    const size_t idx_bytes = (size_t)cells * 4;
    auto table_layout = [&](int64_t nE, int64_t nH) {
        M.coef_kind = 1;
        M.nE = nE;
        M.nH = nH;
        const size_t etab_bytes = (size_t)M.nE * 12 * sizeof(T);
        const size_t htab_bytes = (size_t)M.nH * 20 * sizeof(T);
        M.coef_bytes = 2 * align256(idx_bytes) + align256(etab_bytes) + align256(htab_bytes);
        HIP_OK(hipMalloc(&M.d_coef, M.coef_bytes));
        char* b = (char*)M.d_coef;
        M.eidx = (const uint32_t*)b;
        M.hidx = (const uint32_t*)(b + align256(idx_bytes));
        M.etab = b + 2 * align256(idx_bytes);
        M.htab = b + 2 * align256(idx_bytes) + align256(etab_bytes);
        return std::make_pair(etab_bytes, htab_bytes);
    };
    TuplePtrs<U, 12> ea;
    TuplePtrs<U, 18> ha;
    for (int k = 0; k < 12; ++k) ea.p[k] = (const U*)M.raw[k];
    for (int k = 0; k < 18; ++k) ha.p[k] = (const U*)M.raw[12 + k];
    uint32_t *eix = nullptr, *hix = nullptr;
    U *etb = nullptr, *htb = nullptr;
    auto release = [&] {
        for (void* q : {(void*)eix, (void*)hix, (void*)etb, (void*)htb})
            if (q) (void)hipFree(q);
    };
    const int64_t nE = device_dedup<U, 12>(ea, (uint32_t)cells, limit, 12, &eix, &etb);
    int64_t nH = 0;
    if (nE <= limit) nH = device_dedup<U, 18>(ha, (uint32_t)cells, limit, 20, &hix, &htb);
    if (nE > limit || nH > limit) {
        release();
        if (flags & RFDTD_COEF_TABLE) RF_THROW(RFDTD_EINVAL, "coefficient table exceeds its entry limit");
        return;  // the raw arrays stay
    }
    void* raw_buf = M.d_coef;
    const size_t raw_bytes = M.coef_bytes;
    try {
        M.d_coef = nullptr;
        const auto tb = table_layout(nE, nH);
        HIP_OK(hipMemcpy((void*)M.eidx, eix, idx_bytes, hipMemcpyDeviceToDevice));
        HIP_OK(hipMemcpy((void*)M.hidx, hix, idx_bytes, hipMemcpyDeviceToDevice));
        HIP_OK(hipMemcpy((void*)M.etab, etb, tb.first, hipMemcpyDeviceToDevice));
        HIP_OK(hipMemcpy((void*)M.htab, htb, tb.second, hipMemcpyDeviceToDevice));
    } catch (...) {  // back to the raw form, which the member's destructor frees
        release();
        if (M.d_coef) (void)hipFree(M.d_coef);
        M.d_coef = raw_buf;
        M.coef_bytes = raw_bytes;
        M.coef_kind = 0;
        M.eidx = M.hidx = nullptr;
        M.etab = M.htab = nullptr;
        throw;
    }
    release();
    HIP_OK(hipFree(raw_buf));
    for (int k = 0; k < 30; ++k) M.raw[k] = nullptr;
}

template <typename T>
void upload_coefficients(Member& M, uint32_t flags, const void* const coeff[30]) {
    const int64_t cells = M.cells, plane = (int64_t)M.Ny * M.Nz;
    using U = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    const U* arr[30];
    bool zl[30];
    for (int k = 0; k < 30; ++k) {
        if (!coeff[k]) RF_THROW(RFDTD_EINVAL, "missing coefficient array " + std::to_string(k));
        arr[k] = (const U*)coeff[k];
        zl[k] = zero_last_plane(k);
    }
    // A table is worth it while its entries cost well under the raw arrays.
    const int64_t limit = (flags & RFDTD_COEF_TABLE) ? cells : std::max<int64_t>(4096, cells / 8);
    const size_t idx_bytes = (size_t)cells * 4;
    auto table_layout = [&](int64_t nE, int64_t nH) {
        M.coef_kind = 1;
        M.nE = nE;
        M.nH = nH;
        const size_t etab_bytes = (size_t)M.nE * 12 * sizeof(T);
        const size_t htab_bytes = (size_t)M.nH * 20 * sizeof(T);
        M.coef_bytes = 2 * align256(idx_bytes) + align256(etab_bytes) + align256(htab_bytes);
        HIP_OK(hipMalloc(&M.d_coef, M.coef_bytes));
        char* b = (char*)M.d_coef;
        M.eidx = (const uint32_t*)b;
        M.hidx = (const uint32_t*)(b + align256(idx_bytes));
        M.etab = b + 2 * align256(idx_bytes);
        M.htab = b + 2 * align256(idx_bytes) + align256(etab_bytes);
        return std::make_pair(etab_bytes, htab_bytes);
    };
    if ((flags & RFDTD_HOST_TABLE) && !(flags & RFDTD_COEF_RAW)) {
        std::vector<uint32_t> eidx, hidx;
        std::vector<U> etab, htab;
        const bool use_table = dedup<U, 12>(arr, zl, cells, plane, limit, eidx, etab) &&
                               dedup<U, 18>(arr + 12, zl + 12, cells, plane, limit, hidx, htab);
        if (!use_table && (flags & RFDTD_COEF_TABLE))
            RF_THROW(RFDTD_EINVAL, "coefficient table exceeds its entry limit");
        if (use_table) {
            const auto tb = table_layout((int64_t)etab.size() / 12, (int64_t)htab.size() / 18);
            HIP_OK(hipMemcpy((void*)M.eidx, eidx.data(), idx_bytes, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy((void*)M.hidx, hidx.data(), idx_bytes, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy((void*)M.etab, etab.data(), tb.first, hipMemcpyHostToDevice));
            std::vector<U> hpad((size_t)M.nH * 20, U(0));
            for (int64_t e = 0; e < M.nH; ++e)
                std::memcpy(&hpad[(size_t)e * 20], &htab[(size_t)e * 18], 18 * sizeof(U));
            HIP_OK(hipMemcpy((void*)M.htab, hpad.data(), tb.second, hipMemcpyHostToDevice));
            return;
        }
    }
    // the raw arrays on the device (the last x plane zeroed where the rule applies)
    M.coef_kind = 0;
    const size_t arr_bytes = align256((size_t)cells * sizeof(T));
    M.coef_bytes = 30 * arr_bytes;
    HIP_OK(hipMalloc(&M.d_coef, M.coef_bytes));
    for (int k = 0; k < 30; ++k) {
        char* dst = (char*)M.d_coef + (size_t)k * arr_bytes;
        M.raw[k] = dst;
        HIP_OK(hipMemcpy(dst, coeff[k], (size_t)cells * sizeof(T), hipMemcpyHostToDevice));
        if (zl[k]) HIP_OK(hipMemset(dst + (size_t)(cells - plane) * sizeof(T), 0, (size_t)plane * sizeof(T)));
    }
    if (flags & (RFDTD_COEF_RAW | RFDTD_HOST_TABLE)) return;
    table_from_raw<T>(M, flags);
}

// The member's 30 coefficient arrays computed on the device from the assembly description, then
// deduplicated into the table unless RFDTD_COEF_RAW.
template <typename T>
void assemble_coefficients(Member& M, uint32_t flags, const rfdtd_assembly& a) {
    const int64_t cells = M.cells;
    const int N[3] = {M.Nx, M.Ny, M.Nz};
    if (a.Nx != M.Nx || a.Ny != M.Ny || a.Nz != M.Nz) RF_THROW(RFDTD_EINVAL, "assembly sizes differ from the grid");
    if (!a.eps || !a.sigma) RF_THROW(RFDTD_EINVAL, "cell permittivity and conductivity are required");
    for (int d = 0; d < 3; ++d)
        if (!a.width[d] || !a.inv_dual[d] || !a.inv_width[d]) RF_THROW(RFDTD_EINVAL, "cell widths and inverses are required");
    if (a.n_pml < 0 || a.n_pml > kMaxPmlSides || (a.n_pml && !a.pml)) RF_THROW(RFDTD_EINVAL, "at most 6 PML sides");
    for (int m = 0; m < a.n_pml; ++m) {
        const rfdtd_pml_side& P = a.pml[m];
        if (P.axis < 0 || P.axis > 2 || P.n < 1 || P.n > N[P.axis] - 1 || !P.sigma_edge || !P.sigma_half)
            RF_THROW(RFDTD_EINVAL, "invalid PML side");
    }
    if (a.n_fix < 0 || (a.n_fix && !a.fix) || a.n_scale < 0 || (a.n_scale && !a.scale))
        RF_THROW(RFDTD_EINVAL, "invalid override lists");
    for (int m = 0; m < a.n_fix; ++m) {
        const rfdtd_edge_fix& F = a.fix[m];
        if (F.component < 0 || F.component > 2 || F.kind < 0 || F.kind > 1 || F.count < 0 ||
            (F.count && (!F.cells || !F.values)))
            RF_THROW(RFDTD_EINVAL, "invalid edge override group");
        for (int64_t t = 0; t < F.count; ++t)
            if (F.cells[t] < 0 || F.cells[t] >= cells) RF_THROW(RFDTD_EINVAL, "edge override outside the grid");
    }
    for (int m = 0; m < a.n_scale; ++m) {
        const int c = a.scale[m].coef;
        const bool db = (c >= 14 && c <= 17) || (c >= 20 && c <= 23) || (c >= 26 && c <= 29);
        if (!db || a.scale[m].cell < 0 || a.scale[m].cell >= cells) RF_THROW(RFDTD_EINVAL, "invalid edge-correction factor");
    }
    DedupScratch S;
    auto up = [&](const void* src, size_t bytes) {
        void* d = S.alloc(bytes);
        HIP_OK(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
        return d;
    };
    AsmDev<T> A{};
    A.Nx = M.Nx; A.Ny = M.Ny; A.Nz = M.Nz; A.npml = a.n_pml;
    A.eps = (const double*)up(a.eps, (size_t)cells * 8);
    A.sig = (const double*)up(a.sigma, (size_t)cells * 8);
    for (int d = 0; d < 3; ++d) {
        A.w[d] = (const double*)up(a.width[d], (size_t)N[d] * 8);
        A.idual[d] = (const T*)up(a.inv_dual[d], (size_t)N[d] * sizeof(T));
        A.iw[d] = (const T*)up(a.inv_width[d], (size_t)N[d] * sizeof(T));
    }
    A.dt = a.dt; A.e0 = a.e0; A.u0 = a.u0;
    for (int m = 0; m < a.n_pml; ++m) {
        const rfdtd_pml_side& P = a.pml[m];
        A.pml[m] = AsmPml{P.axis, P.upper ? 1 : 0, P.n, 0, (const double*)up(P.sigma_edge, (size_t)P.n * 8),
                          (const double*)up(P.sigma_half, (size_t)P.n * 8)};
    }
    // the raw arrays, then the passes
    M.coef_kind = 0;
    const size_t arr_bytes = align256((size_t)cells * sizeof(T));
    M.coef_bytes = 30 * arr_bytes;
    HIP_OK(hipMalloc(&M.d_coef, M.coef_bytes));
    T* rawp[30];
    for (int k = 0; k < 30; ++k) {
        rawp[k] = (T*)((char*)M.d_coef + (size_t)k * arr_bytes);
        M.raw[k] = rawp[k];
    }
    T* const* d_raw = (T* const*)up(rawp, sizeof(rawp));
    const int blocks = (int)((cells + 255) / 256);
    k_asm_dense<T><<<blocks, 256>>>(A, d_raw);
    HIP_OK(hipGetLastError());
    for (int m = 0; m < a.n_fix; ++m) {
        const rfdtd_edge_fix& F = a.fix[m];
        if (!F.count) continue;
        const int64_t* dc = (const int64_t*)up(F.cells, (size_t)F.count * 8);
        const double* dv = (const double*)up(F.values, (size_t)(F.kind == 0 ? 1 : F.count) * 8);
        k_asm_fix<T><<<(int)((F.count + 255) / 256), 256>>>(A, d_raw, F.component, F.kind, F.wide ? 1 : 0, F.count, dc, dv);
        HIP_OK(hipGetLastError());
    }
    if (a.n_scale) {
        const rfdtd_h_scale* ds = (const rfdtd_h_scale*)up(a.scale, (size_t)a.n_scale * sizeof(rfdtd_h_scale));
        k_asm_scale<T><<<1, 64>>>(d_raw, a.n_scale, ds);
        HIP_OK(hipGetLastError());
    }
    k_asm_pack<T><<<blocks, 256>>>(A, d_raw);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    if (!(flags & RFDTD_COEF_RAW)) table_from_raw<T>(M, flags);
}

// x planes per workgroup: enough workgroups to give every CU several waves (>= ~2048 groups in a
// launch on 256 CUs), as many planes as possible for register reuse along x.
// Workgroups per x plane of a member: (y,z) columns per workgroup, or fused tiles.
int64_t tiles_per_plane(const Session& s, const Member& M) {
    if (s.two_step) return (int64_t)((M.Ny + s.fused - 1) / s.fused) * ((M.Nz + s.p2_rs - 5) / (s.p2_rs - 4));
    if (s.fused) return (int64_t)((M.Ny + s.fused - 1) / s.fused) * ((M.Nz + kFz - 1) / kFz);
    return ((int64_t)M.Ny * M.Nz / s.vec + s.block - 1) / s.block;
}

// Cells per thread of the separate kernels for a variant (1, or 2 / 4 for the vector columns), and
// whether every grid of the session allows it (Nz a multiple of it).
int variant_vec(int v) { return v == 11 ? 2 : v == 12 ? 4 : 1; }
bool vec_fits(const Session& s, int vec) {
    for (const auto& M : s.members)
        if (M.Nz % vec) return false;
    return true;
}

// The session's kernel flags for a variant: 0 separate kernels, 1 separate pipelined, 2 fused 6-row
// tiles, 3 fused 14-row tiles, 4 = 1 with nontemporal stores, 5 and 6 = 2 and 3 with E one plane ahead
// of H, 7 and 8 the two-step pass with 12- and 4-row tiles, 9 its one-barrier form with 8-row tiles,
// 10 = 4 with nontemporal loads, 11 and 12 = 10 over 2 and 4 consecutive z cells per thread, 13 and
// 14 the two-step pass with 32-lane rows (28 x 28 and 12 x 28 cell tiles).
bool is_two_step(int v) { return v == 7 || v == 8 || v == 9 || v == 13 || v == 14 || v == 15; }
void apply_variant(Session& s, int v) {
    s.pipelined = v != 0;
    s.two_step = is_two_step(v);
    s.two_skew = v == 9;
    s.p2_rs = (v == 13 || v == 14 || v == 15) ? 32 : 64;
    s.fused = v == 7 ? kP2TyTall : v == 8 ? kP2TyShort : v == 9 ? kP2TySkew : v == 13 ? kP2Ty32Tall
            : v == 14 ? kP2Ty32Short : v == 15 ? kP2Ty32Mid : (v == 2 || v == 5) ? 6 : (v == 3 || v == 6) ? 14 : 0;
    s.fused_deep = v == 5 || v == 6;
    // the separate kernels' load and store hints (the tiled passes have their own)
    s.nontemporal = v == 4 || (v >= 10 && v <= 12);
    s.nt_loads = v >= 10 && v <= 12;
    s.vec = variant_vec(v);
    s.dirty = true;
    s.graph_dirty = true;
}

int variant_of(const Session& s) {
    return s.vec == 4 ? 12 : s.vec == 2 ? 11 : s.nt_loads ? 10 : s.two_skew ? 9
         : (s.two_step && s.p2_rs == 32) ? (s.fused == kP2Ty32Tall ? 13 : s.fused == kP2Ty32Mid ? 15 : 14)
         : s.two_step ? (s.fused == kP2TyShort ? 8 : 7) : s.fused == 14 ? (s.fused_deep ? 6 : 3)
         : s.fused == 6 ? (s.fused_deep ? 5 : 2) : s.nontemporal ? 4 : s.pipelined ? 1 : 0;
}

void plan_launch(Session& s) {
    int64_t blocks_per_plane = 0;
    for (auto& M : s.members) blocks_per_plane += M.K * tiles_per_plane(s, M);
    // workgroups to aim for: the two-step pass holds one workgroup per CU, and about three rounds of
    // them (67-plane chunks on 400x400x200) ran faster than eight
    const int64_t want = s.two_step ? 768 : 2048;
    for (auto& M : s.members) {
        if (M.xplanes_req > 0) {
            M.xplanes = std::min(M.xplanes_req, M.Nx);
            continue;
        }
        int64_t chunks = std::max<int64_t>(1, (want + blocks_per_plane - 1) / blocks_per_plane);
        chunks = std::min<int64_t>(chunks, M.Nx);
        M.xplanes = (int)((M.Nx + chunks - 1) / chunks);
    }
}

// Block offsets of a batched flat launch over items of the given sizes (256 threads per block).
std::vector<int64_t> block_offsets(const std::vector<int64_t>& sizes, int64_t block = 256) {
    std::vector<int64_t> off(sizes.size() + 1, 0);
    for (size_t i = 0; i < sizes.size(); ++i) off[i + 1] = off[i] + (sizes[i] + block - 1) / block;
    if (off.back() > INT32_MAX) RF_THROW(RFDTD_EINVAL, "batched launch too large");
    return off;
}

// Probe groups of the instance sorted by cell, E and H apart, with a bitmask of their cells: the
// update kernels (and the fused pass, for E) apply them where the values are computed.
template <typename T>
void build_probe_lookup(Instance& I, const Member& M) {
    const int nG = (int)I.groups.size();
    for (int h = 0; h < 2; ++h) {
        const int g0 = h ? I.ngE : 0, g1 = h ? nG : I.ngE;
        uint32_t*& d_bits = h ? I.d_hbits : I.d_ebits;
        void*& d_groups = h ? I.d_hgroups : I.d_egroups;
        int& count = h ? I.nhgroups : I.negroups;
        if (d_bits || g1 <= g0) continue;
        std::vector<ProbeGroup<T>> g;
        for (int k = g0; k < g1; ++k) {
            const auto& hg = I.groups[k];
            ProbeGroup<T> pg{};
            pg.cell = hg.cell;
            pg.series = (T*)I.d_series;
            pg.series_in = (const T*)I.d_series_in;
            pg.order = I.d_order + hg.start;
            pg.is_src = I.d_is_src;
            pg.field = hg.field;
            pg.count = hg.count;
            pg.nprobes = I.P;
            g.push_back(pg);
        }
        std::stable_sort(g.begin(), g.end(),
                         [](const ProbeGroup<T>& a, const ProbeGroup<T>& b) { return a.cell < b.cell; });
        std::vector<uint32_t> bits((size_t)(M.cells + 31) / 32, 0u);
        for (const auto& pg : g) bits[(size_t)(pg.cell >> 5)] |= 1u << (pg.cell & 31);
        HIP_OK(hipMalloc(&d_bits, bits.size() * 4));
        HIP_OK(hipMemcpy(d_bits, bits.data(), bits.size() * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMalloc(&d_groups, g.size() * sizeof(ProbeGroup<T>)));
        HIP_OK(hipMemcpy(d_groups, g.data(), g.size() * sizeof(ProbeGroup<T>), hipMemcpyHostToDevice));
        count = (int)g.size();
    }
}

// Kernel descriptor of a field buffer (18 arrays at the given pitch) on member M's grid.
template <typename T>
InstDev<T> make_instdev(const Member& M, void* fields, size_t pitch, int npb, void* out = nullptr) {
    InstDev<T> d;
    std::memset(&d, 0, sizeof(d));
    T** a = (T**)&d.F;
    for (int k = 0; k < 18; ++k) a[k] = (T*)((char*)fields + (size_t)k * pitch);
    if (out) {
        T** g = (T**)&d.G;
        for (int k = 0; k < 18; ++k) g[k] = (T*)((char*)out + (size_t)k * pitch);
    }
    for (int k = 0; k < 30; ++k) d.raw[k] = (const T*)M.raw[k];
    d.eidx = M.eidx;
    d.hidx = M.hidx;
    d.etab = (const Vec4<T>*)M.etab;
    d.htab = (const Vec4<T>*)M.htab;
    d.coef_kind = M.coef_kind;
    d.Nx = M.Nx; d.Ny = M.Ny; d.Nz = M.Nz;
    d.xplanes = M.xplanes;
    d.npb = npb;
    d.tz_n = (M.Nz + kFz - 1) / kFz;
    return d;
}

template <typename T>
void refresh(Session& s) {
    if (!s.dirty) return;
    plan_launch(s);
    if (s.fused)
        for (Instance& I : s.insts)
            if (!I.d_fields2) {
                HIP_OK(hipMalloc(&I.d_fields2, I.fields_bytes));
                HIP_OK(hipMemset(I.d_fields2, 0, I.fields_bytes));
            }
    // instance descriptors: [parity 0] reads buffer 1 of each pair and writes buffer 2, [parity 1]
    // the other way round (the separate kernels only use F, the current buffer)
    std::vector<InstDev<T>> di(s.insts.size()), di2(s.insts.size());
    std::vector<MemberDev> dm(s.members.size());
    int tile0 = 0;
    for (size_t m = 0; m < s.members.size(); ++m) {
        Member& M = s.members[m];
        const int npb = (int)tiles_per_plane(s, M);
        const int chunks = (M.Nx + M.xplanes - 1) / M.xplanes;
        dm[m] = MemberDev{tile0, npb * chunks, M.K, M.inst0, s.col_order};
        tile0 += npb * chunks * M.K;
        for (int e = 0; e < M.K; ++e) {
            Instance& I = s.insts[M.inst0 + e];
            build_probe_lookup<T>(I, M);
            di[M.inst0 + e] = make_instdev<T>(M, I.d_fields, I.pitch, npb, I.d_fields2);
            di2[M.inst0 + e] = I.d_fields2 ? make_instdev<T>(M, I.d_fields2, I.pitch, npb, I.d_fields) : di[M.inst0 + e];
            for (InstDev<T>* d : {&di[M.inst0 + e], &di2[M.inst0 + e]}) {
                d->ebits = I.d_ebits;
                d->egroups = I.d_egroups;
                d->negroups = I.negroups;
                d->hbits = I.d_hbits;
                d->hgroups = I.d_hgroups;
                d->nhgroups = I.nhgroups;
            }
        }
    }
    s.total_tiles = tile0;
    for (void** p : {&s.d_insts, &s.d_insts2, (void**)&s.d_members, (void**)&s.d_flat0}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    HIP_OK(hipMalloc(&s.d_insts, di.size() * sizeof(InstDev<T>)));
    HIP_OK(hipMemcpy(s.d_insts, di.data(), di.size() * sizeof(InstDev<T>), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&s.d_insts2, di2.size() * sizeof(InstDev<T>)));
    HIP_OK(hipMemcpy(s.d_insts2, di2.data(), di2.size() * sizeof(InstDev<T>), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&s.d_members, dm.size() * sizeof(MemberDev)));
    HIP_OK(hipMemcpy(s.d_members, dm.data(), dm.size() * sizeof(MemberDev), hipMemcpyHostToDevice));
    std::vector<int64_t> cells;
    for (const Instance& I : s.insts) cells.push_back(s.members[I.member].cells);
    s.flat0 = block_offsets(cells);
    HIP_OK(hipMalloc(&s.d_flat0, s.flat0.size() * sizeof(int64_t)));
    HIP_OK(hipMemcpy(s.d_flat0, s.flat0.data(), s.flat0.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    for (void** p : {&s.d_hist_boxes, (void**)&s.d_hist_h0, (void**)&s.d_hist_hb}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    s.hist_blocks = 0;
    s.hist_blocks_e = 0;
    if (!s.hist.empty()) {
        std::vector<BoxDev<T>> hb(s.hist.size());
        std::vector<int64_t> hc;
        for (size_t k = 0; k < s.hist.size(); ++k) {
            const auto& h = s.hist[k];
            hb[k].B = Box{h.box[0], h.box[2], h.box[4], h.box[1] - h.box[0], h.box[3] - h.box[2], h.box[5] - h.box[4]};
            hb[k].inst = h.inst;
            hb[k].slots = (T*)h.d;
            hb[k].grad = nullptr;
            hb[k].slot_elems = h.slot_elems;
            hc.push_back(h.slot_elems / 3);
        }
        const std::vector<int64_t> h0 = block_offsets(hc);
        s.hist_blocks = (int)h0.back();
        HIP_OK(hipMalloc(&s.d_hist_boxes, hb.size() * sizeof(BoxDev<T>)));
        HIP_OK(hipMemcpy(s.d_hist_boxes, hb.data(), hb.size() * sizeof(BoxDev<T>), hipMemcpyHostToDevice));
        HIP_OK(hipMalloc(&s.d_hist_h0, h0.size() * sizeof(int64_t)));
        HIP_OK(hipMemcpy(s.d_hist_h0, h0.data(), h0.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        const std::vector<int64_t> he = block_offsets(hc, s.block);
        s.hist_blocks_e = (int)he.back();
        HIP_OK(hipMalloc(&s.d_hist_hb, he.size() * sizeof(int64_t)));
        HIP_OK(hipMemcpy(s.d_hist_hb, he.data(), he.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    }

    // probe groups of all instances: E groups, then H groups
    std::vector<ProbeGroup<T>> ge, gh;
    for (size_t k = 0; k < s.insts.size(); ++k) {
        const Instance& I = s.insts[k];
        for (size_t g = 0; g < I.groups.size(); ++g) {
            const auto& hg = I.groups[g];
            ProbeGroup<T> pg{};
            pg.cell = hg.cell;
            pg.series = (T*)I.d_series;
            pg.series_in = (const T*)I.d_series_in;
            pg.order = I.d_order + hg.start;
            pg.is_src = I.d_is_src;
            pg.inst = (int32_t)k;
            pg.field = hg.field;
            pg.count = hg.count;
            pg.nprobes = I.P;
            ((int)g < I.ngE ? ge : gh).push_back(pg);
        }
    }
    s.ngE = (int)ge.size();
    s.ngH = (int)gh.size();
    ge.insert(ge.end(), gh.begin(), gh.end());
    if (s.d_groups) (void)hipFree(s.d_groups);
    s.d_groups = nullptr;
    if (!ge.empty()) {
        HIP_OK(hipMalloc(&s.d_groups, ge.size() * sizeof(ProbeGroup<T>)));
        HIP_OK(hipMemcpy(s.d_groups, ge.data(), ge.size() * sizeof(ProbeGroup<T>), hipMemcpyHostToDevice));
    }

    // monitors of all instances
    std::vector<MonitorDev> md;
    int blocks = 0, fblocks = 0;
    size_t stage_elems = 0;
    for (size_t k = 0; k < s.insts.size(); ++k)
        for (const MonitorHost& h : s.insts[k].mons) {
            if (!h.active) continue;
            MonitorDev d{};
            d.inst = (int32_t)k;
            d.id = (int32_t)md.size();
            d.axis = h.desc.axis; d.pos = h.pos; d.field = h.desc.field; d.n_step = h.desc.n_step;
            d.nf = h.desc.n_frequencies; d.N1 = h.N1; d.N2 = h.N2;
            d.ring = h.ring;
            d.acc64 = (h.desc.flags & RFDTD_MONITOR_DFT64) && !s.f64;
            d.N1N2 = (int64_t)h.N1 * h.N2;
            d.blk0 = blocks;
            d.values = h.d_values;
            d.phase = h.d_phase;
            d.stage = (void*)stage_elems;  // offset for now
            if (d.nf) stage_elems += (size_t)kStageSlots * d.N1N2;
            d.fblk0 = fblocks;
            const int G = (d.acc64 || s.f64) ? flush_group<double>() : flush_group<float>();
            fblocks += d.nf ? (int)((d.N1N2 + 255) / 256) * ((d.nf + G - 1) / G) : 0;
            blocks += (int)((d.N1N2 + 255) / 256);
            md.push_back(d);
        }
    s.n_active = (int)md.size();
    s.mon_blocks = blocks;
    s.flush_blocks = fblocks;
    if (s.d_mons) (void)hipFree(s.d_mons);
    if (s.d_stage) (void)hipFree(s.d_stage);
    if (s.d_rows) (void)hipFree(s.d_rows);
    s.d_mons = nullptr;
    s.d_stage = nullptr;
    s.d_rows = nullptr;
    if (stage_elems) HIP_OK(hipMalloc(&s.d_stage, stage_elems * sizeof(T)));
    for (auto& d : md) d.stage = d.nf ? (void*)((T*)s.d_stage + (size_t)d.stage) : nullptr;
    if (!md.empty()) {
        HIP_OK(hipMalloc(&s.d_rows, md.size() * kStageSlots * sizeof(int)));
        HIP_OK(hipMemset(s.d_rows, 0xff, md.size() * kStageSlots * sizeof(int)));
        HIP_OK(hipMalloc(&s.d_mons, md.size() * sizeof(MonitorDev)));
        HIP_OK(hipMemcpy(s.d_mons, md.data(), md.size() * sizeof(MonitorDev), hipMemcpyHostToDevice));
    }
    s.dirty = false;
    s.drop_graph();
}

// Pairs of steps go through the two-step pass when its variant is chosen, no design-box history is
// being recorded (it needs H at every step) and the pass can sample every active monitor itself.
bool pairs_ok(const Session& s) { return s.two_step && s.hist.empty() && s.n_active <= kP2Mon; }

// The two-step pass over every instance (two: steps local and local + 1, else step local alone),
// from the current buffer of each pair into the other.
template <typename T>
void launch_step2(Session& s, hipStream_t st, const int* base, int local, bool two, int slot, hipEvent_t* ev) {
    const int grid = (s.total_tiles + 7) & ~7;
    const InstDev<T>* di = (const InstDev<T>*)(s.parity ? s.d_insts2 : s.d_insts);
    const int io = s.tuning ? 0 : 1;
    if (ev) HIP_OK(hipEventRecord(ev[0], st));
    if (s.two_skew)
        k_step2<T, kP2TySkew, true><<<grid, (kP2TySkew + 4) * 64, 0, st>>>(
            di, s.d_members, (int)s.members.size(), s.total_tiles, base, local, io, two ? 1 : 0, s.d_mons, s.n_active,
            slot, s.d_rows);
    else if (s.p2_rs == 32 && s.fused == kP2Ty32Tall)
        k_step2<T, kP2Ty32Tall, false, 32><<<grid, (kP2Ty32Tall + 4) * 32, 0, st>>>(
            di, s.d_members, (int)s.members.size(), s.total_tiles, base, local, io, two ? 1 : 0, s.d_mons, s.n_active,
            slot, s.d_rows);
    else if (s.p2_rs == 32 && s.fused == kP2Ty32Mid)
        k_step2<T, kP2Ty32Mid, false, 32><<<grid, (kP2Ty32Mid + 4) * 32, 0, st>>>(
            di, s.d_members, (int)s.members.size(), s.total_tiles, base, local, io, two ? 1 : 0, s.d_mons, s.n_active,
            slot, s.d_rows);
    else if (s.p2_rs == 32)
        k_step2<T, kP2Ty32Short, false, 32><<<grid, (kP2Ty32Short + 4) * 32, 0, st>>>(
            di, s.d_members, (int)s.members.size(), s.total_tiles, base, local, io, two ? 1 : 0, s.d_mons, s.n_active,
            slot, s.d_rows);
    else if (s.fused == kP2TyShort)
        k_step2<T, kP2TyShort, false><<<grid, (kP2TyShort + 4) * 64, 0, st>>>(
            di, s.d_members, (int)s.members.size(), s.total_tiles, base, local, io, two ? 1 : 0, s.d_mons, s.n_active,
            slot, s.d_rows);
    else
        k_step2<T, kP2TyTall, false><<<grid, (kP2TyTall + 4) * 64, 0, st>>>(
            di, s.d_members, (int)s.members.size(), s.total_tiles, base, local, io, two ? 1 : 0, s.d_mons, s.n_active,
            slot, s.d_rows);
    if (ev) {
        HIP_OK(hipEventRecord(ev[1], st));
        HIP_OK(hipEventRecord(ev[2], st));
        HIP_OK(hipEventRecord(ev[3], st));
    }
    s.parity ^= 1;
    s.totals_stale = true;
}

// Steps local and local + 1 in one two-step pass.  slot_a, slot_b: their staging slots; the pass
// samples step local's monitors itself, k_monitors step local + 1's from the stored splits; flush > 0
// accumulates staged slots [0, flush) afterwards.
template <typename T>
void launch_pair(Session& s, hipStream_t st, const int* base, int local, int slot_a, int slot_b, int flush,
                 hipEvent_t* ev = nullptr) {
    launch_step2<T>(s, st, base, local, true, slot_a, ev);
    if (!s.tuning && s.n_active) {
        const InstDev<T>* di = (const InstDev<T>*)(s.parity ? s.d_insts2 : s.d_insts);
        k_monitors<T><<<s.mon_blocks, 256, 0, st>>>(di, s.d_mons, s.n_active, base, local + 1, slot_b, s.d_rows, 0);
        if (flush && s.flush_blocks)
            k_monitor_flush<T><<<(s.flush_blocks + 7) & ~7, 256, 0, st>>>(s.d_mons, s.n_active, s.flush_blocks, flush,
                                                                       s.d_rows);
    }
}

// One time step.  slot: the step's staging slot for running DFTs; flush > 0 accumulates staged
// slots [0, flush) after this step's sample.
template <typename T>
void launch_step(Session& s, hipStream_t st, const int* base, int local, int slot, int flush,
                 hipEvent_t* ev = nullptr) {
    const int grid = (s.total_tiles + 7) & ~7;
    const InstDev<T>* desc[2] = {(const InstDev<T>*)s.d_insts, (const InstDev<T>*)s.d_insts2};
    const InstDev<T>* di = desc[s.parity];
    const int nmem = (int)s.members.size();
    const bool io = !s.tuning;
    // H^n of the design boxes before this step's E half step: in the E launch for the separate
    // kernels, a launch of its own before the fused pass
    const bool save_in_e = io && s.hist_blocks && !s.fused && s.vec == 1;
    if (io && s.hist_blocks && (s.fused || s.vec > 1))
        k_save_h<T><<<s.hist_blocks, 256, 0, st>>>(di, (const BoxDev<T>*)s.d_hist_boxes, s.d_hist_h0, (int)s.hist.size(),
                                                  local, base);
    if (s.two_step) {
        launch_step2<T>(s, st, base, local, false, slot, ev);  // one step through the two-step pass
        di = desc[s.parity];
    } else if (s.fused) {
        // one pass writes the other buffer of each pair; monitors then sample it
        if (ev) HIP_OK(hipEventRecord(ev[0], st));
        if (s.fused == 14 && s.fused_deep)
            k_step_fused<T, 14, true><<<grid, 16 * 64, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, io);
        else if (s.fused == 14)
            k_step_fused<T, 14, false><<<grid, 16 * 64, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, io);
        else if (s.fused_deep)
            k_step_fused<T, 6, true><<<grid, 8 * 64, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, io);
        else
            k_step_fused<T, 6, false><<<grid, 8 * 64, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, io);
        if (ev) {
            HIP_OK(hipEventRecord(ev[1], st));
            HIP_OK(hipEventRecord(ev[2], st));
            HIP_OK(hipEventRecord(ev[3], st));
        }
        s.parity ^= 1;
        s.totals_stale = true;
        di = desc[s.parity];
    } else if (s.vec > 1) {
        // the separate kernels over 2 or 4 z cells per thread, probes applied inside
        const int pio = io ? 1 : 0;
        if (ev) HIP_OK(hipEventRecord(ev[0], st));
        if (s.vec == 2)
            k_update_e_w<T, 2, true><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        else
            k_update_e_w<T, 4, false><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        if (ev) {
            HIP_OK(hipEventRecord(ev[1], st));
            HIP_OK(hipEventRecord(ev[2], st));
        }
        if (s.vec == 2)
            k_update_h_w<T, 2, true><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        else
            k_update_h_w<T, 4, false><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        if (ev) HIP_OK(hipEventRecord(ev[3], st));
    } else {
        // E and H probes (sources and records) are applied inside the update kernels
        const int pio = io ? 1 : 0;
        if (ev) HIP_OK(hipEventRecord(ev[0], st));
        if (save_in_e) {
            const int gs = (s.total_tiles + s.hist_blocks_e + 7) & ~7;
            const BoxDev<T>* hb = (const BoxDev<T>*)s.d_hist_boxes;
            const int nb = (int)s.hist.size();
            if (s.nontemporal)
                k_update_e_save<T, true, true><<<gs, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local,
                                                                      pio, hb, s.d_hist_hb, nb);
            else if (s.pipelined)
                k_update_e_save<T, true, false><<<gs, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local,
                                                                       pio, hb, s.d_hist_hb, nb);
            else
                k_update_e_save<T, false, false><<<gs, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base,
                                                                        local, pio, hb, s.d_hist_hb, nb);
        } else if (s.nt_loads)
            k_update_e<T, true, true, true><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local,
                                                                     pio);
        else if (s.nontemporal)
            k_update_e<T, true, true><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        else if (s.pipelined)
            k_update_e<T, true><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        else
            k_update_e<T, false><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        if (ev) {
            HIP_OK(hipEventRecord(ev[1], st));
            HIP_OK(hipEventRecord(ev[2], st));
        }
        if (s.nt_loads)
            k_update_h<T, true, true, true><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local,
                                                                     pio);
        else if (s.nontemporal)
            k_update_h<T, true, true><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        else if (s.pipelined)
            k_update_h<T, true><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        else
            k_update_h<T, false><<<grid, s.block, 0, st>>>(di, s.d_members, nmem, s.total_tiles, base, local, pio);
        if (ev) HIP_OK(hipEventRecord(ev[3], st));
    }

    if (io && s.n_active) {
        k_monitors<T><<<s.mon_blocks, 256, 0, st>>>(di, s.d_mons, s.n_active, base, local, slot, s.d_rows,
                                                    s.fused ? 0 : 1);
        if (flush && s.flush_blocks)
            k_monitor_flush<T><<<(s.flush_blocks + 7) & ~7, 256, 0, st>>>(s.d_mons, s.n_active, s.flush_blocks, flush,
                                                                       s.d_rows);
    }
}

// After fused steps the separate kernels need the totals of the current buffer again.
template <typename T>
void rebuild_totals(Session& s, hipStream_t st) {
    if (!s.totals_stale) return;
    const InstDev<T>* di = s.parity ? (const InstDev<T>*)s.d_insts2 : (const InstDev<T>*)s.d_insts;
    k_totals<T><<<(int)s.flat0.back(), 256, 0, st>>>(di, s.d_flat0, (int)s.insts.size());
    HIP_OK(hipGetLastError());
    s.totals_stale = false;
}

void save_checkpoints(Session& s, int slot, hipStream_t st) {
    for (Instance& I : s.insts) {
        if ((int)I.ckpt.size() <= slot) I.ckpt.resize(slot + 1, nullptr);
        if (!I.ckpt[slot]) HIP_OK(hipMalloc(&I.ckpt[slot], 12 * I.pitch));
        HIP_OK(hipMemcpyAsync(I.ckpt[slot], s.cur(I), 12 * I.pitch, hipMemcpyDeviceToDevice, st));
    }
}

// Steps per captured chunk: even for the fused step, so that a replay ends on the buffer it began on.
// (the two-step pass flips the buffers once per pair: a multiple of 4 steps ends where it began)
int chunk_steps(const Session& s) {
    if (s.two_step) return std::max(4, s.graph_steps & ~3);
    return s.fused ? std::max(2, s.graph_steps & ~1) : s.graph_steps;
}

// Capture one chunk of steps starting from the current parity (the graph for that parity).
template <typename T>
void capture_graph(Session& s) {
    if (s.graph_dirty) s.drop_graph();
    s.graph_dirty = false;
    const int p = s.parity, n = chunk_steps(s);
    hipStream_t cs;
    HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    const bool pairs = pairs_ok(s);
    for (int k = 0; k < n;) {
        const int slot = k % kStageSlots;
        if (pairs && k + 1 < n) {
            const int slot2 = (k + 1) % kStageSlots;
            const bool last = slot2 == kStageSlots - 1 || k + 1 == n - 1;
            launch_pair<T>(s, cs, s.d_step_base, k, slot, slot2, last ? slot2 + 1 : 0);
            k += 2;
        } else {
            const bool last = slot == kStageSlots - 1 || k == n - 1;
            launch_step<T>(s, cs, s.d_step_base, k, slot, last ? slot + 1 : 0);
            k += 1;
        }
    }
    hipError_t ce = hipStreamEndCapture(cs, &s.graph_def[p]);
    (void)hipStreamDestroy(cs);
    s.parity = p;  // an even number of fused steps: the chunk ends where it began
    HIP_OK(ce);
    HIP_OK(hipGraphInstantiate(&s.graph[p], s.graph_def[p], nullptr, nullptr, 0));
}

// len steps starting at s.step + offset (s.step itself is advanced by the caller)
template <typename T>
void run_chunk(Session& s, int len, int offset, hipStream_t st, bool graphs, bool timing, int stride) {
    int done = 0;
    const int gs = chunk_steps(s);
    if (!s.fused) rebuild_totals<T>(s, st);
    if (graphs && len >= gs) {
        if (s.graph_dirty || !s.graph[s.parity]) capture_graph<T>(s);
        while (len - done >= gs) {
            k_set_step<<<1, 1, 0, st>>>(s.d_step_base, (int)(s.step + offset + done));
            HIP_OK(hipGraphLaunch(s.graph[s.parity], st));
            done += gs;
        }
        if (s.fused) s.totals_stale = true;
    }
    const bool pairs = pairs_ok(s);
    for (int d = 0; done < len;) {
        const int64_t n = s.step + offset + done;
        const int slot = d % kStageSlots;
        hipEvent_t* ev = nullptr;
        if (timing && (offset + done) % stride == 0 && s.n_timed < Session::kTimedSteps) {
            if (s.kev.empty()) {
                s.kev.resize(4 * Session::kTimedSteps);
                for (auto& e : s.kev) HIP_OK(hipEventCreate(&e));
            }
            ev = &s.kev[4 * s.n_timed++];
        }
        if (pairs && done + 1 < len) {  // (d is even here: the staging slots of a pair never straddle a flush)
            const int slot2 = (d + 1) % kStageSlots;
            const bool last = slot2 == kStageSlots - 1 || done + 1 == len - 1;
            launch_pair<T>(s, st, nullptr, (int)n, slot, slot2, last ? slot2 + 1 : 0, ev);
            done += 2;
            d += 2;
        } else {
            const bool last = slot == kStageSlots - 1 || done == len - 1;
            launch_step<T>(s, st, nullptr, (int)n, slot, last ? slot + 1 : 0, ev);
            done += 1;
            d += 1;
        }
    }
}

// Streamed monitors: the sample planes that steps [a, b) produced go from their ring slots to the
// caller's host arrays on the copy stream, after the run segment on `st`; the segment two ahead
// waits for these copies before it overwrites the slots.
void stream_monitors(Session& s, hipStream_t st, int64_t a, int64_t b) {
    HIP_OK(hipEventRecord(s.seg_ev, st));
    HIP_OK(hipStreamWaitEvent(s.copy_stream, s.seg_ev, 0));
    auto samples = [](int64_t N, int p) { return N <= 0 ? (int64_t)0 : (N - 1) / p + 1; };
    for (Instance& I : s.insts)
        for (auto& h : I.mons) {
            if (!h.ring) continue;
            const size_t plane = (size_t)h.N1 * h.N2 * s.esz;
            int64_t m = std::min<int64_t>(samples(a, h.desc.n_step), h.Nm);
            const int64_t m1 = std::min<int64_t>(samples(b, h.desc.n_step), h.Nm);
            while (m < m1) {
                const int64_t slot = m % h.ring, cnt = std::min<int64_t>(m1 - m, h.ring - slot);
                HIP_OK(hipMemcpyAsync((char*)h.desc.values + (size_t)m * plane, (const char*)h.d_values + (size_t)slot * plane,
                                      (size_t)cnt * plane, hipMemcpyDeviceToHost, s.copy_stream));
                m += cnt;
            }
        }
    const int e = (int)(s.seg_count & 1);
    HIP_OK(hipEventRecord(s.cpy_ev[e], s.copy_stream));
    s.cpy_pending[e] = true;
    ++s.seg_count;
}

bool has_streamed_monitors(const Session& s) {
    for (const Instance& I : s.insts)
        for (const auto& h : I.mons)
            if (h.ring) return true;
    return false;
}

template <typename T>
void run_steps(Session& s, int n_steps, hipStream_t st) {
    if (n_steps < 0) RF_THROW(RFDTD_EINVAL, "negative step count");
    bool has_probes = false, has_mons = false;
    for (const auto& I : s.insts) {
        has_probes |= I.P > 0;
        for (const auto& m : I.mons) has_mons |= m.active;
    }
    if (has_probes && s.step + n_steps > s.Nt_probes)
        RF_THROW(RFDTD_ESTATE, "run past the end of the probe series (Nt = " + std::to_string(s.Nt_probes) + ")");
    if (has_mons && s.step + n_steps > s.Nt_monitors)
        RF_THROW(RFDTD_ESTATE, "run past the monitor length (Nt = " + std::to_string(s.Nt_monitors) + ")");
    for (const auto& h : s.hist)
        if (s.step + n_steps > h.capacity)
            RF_THROW(RFDTD_ESTATE, "run past the design-box history (" + std::to_string(h.capacity) + " steps)");
    refresh<T>(s);
    if (s.design_pending) {
        HIP_OK(hipStreamWaitEvent(st, s.design_ev, 0));
        s.design_pending = false;
    }
    HIP_OK(hipEventRecord(s.ev0, st));
    const bool timing = (s.flags & RFDTD_KERNEL_TIMING) != 0;
    const bool graphs = !timing && !(s.flags & RFDTD_NO_GRAPHS) && s.graph_steps > 1;
    // kernel timing: events around the E and H updates of up to kTimedSteps evenly spread steps
    s.n_timed = 0;
    const int stride = timing ? std::max(1, n_steps / Session::kTimedSteps) : 0;
    // streamed monitors: segments short enough that each produces at most half a ring of samples
    // (whole graph chunks where the ring allows)
    const bool streaming = has_streamed_monitors(s);
    int64_t seg_len = INT64_MAX;
    if (streaming) {
        for (const Instance& I : s.insts)
            for (const auto& h : I.mons)
                if (h.ring) seg_len = std::min<int64_t>(seg_len, (int64_t)std::max(1, h.ring / 2) * h.desc.n_step);
        const int gs = chunk_steps(s);
        if (graphs && seg_len >= gs) seg_len = seg_len / gs * gs;
        if (!s.copy_stream) {
            HIP_OK(hipStreamCreateWithFlags(&s.copy_stream, hipStreamNonBlocking));
            HIP_OK(hipEventCreateWithFlags(&s.seg_ev, hipEventDisableTiming));
            for (auto& e : s.cpy_ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
    }
    // checkpointed runs save the split fields before every multiple of ckpt_every and replay graph
    // chunks only between checkpoints
    int done = 0;
    while (done < n_steps) {
        const int64_t n = s.step + done;
        int64_t len = n_steps - done;
        if (s.ckpt_every > 0) {
            if (n % s.ckpt_every == 0) save_checkpoints(s, (int)(n / s.ckpt_every), st);
            len = std::min<int64_t>(len, (n / s.ckpt_every + 1) * s.ckpt_every - n);
        }
        if (streaming) {
            len = std::min(len, seg_len);
            const int e = (int)(s.seg_count & 1);  // the copies of the segment two back
            if (s.cpy_pending[e]) HIP_OK(hipStreamWaitEvent(st, s.cpy_ev[e], 0));
        }
        run_chunk<T>(s, (int)len, done, st, graphs, timing, stride);
        if (streaming) stream_monitors(s, st, n, n + len);
        done += (int)len;
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(s.ev1, st));
    s.step += n_steps;
}

// The separate kernels use one field buffer per instance: after a switch away from the fused pass,
// bring the current fields into the first buffer and free the second.
void release_second_buffers(Session& s) {
    bool any = false;
    for (const Instance& I : s.insts) any |= I.d_fields2 != nullptr;
    if (!any) return;
    HIP_OK(hipDeviceSynchronize());
    for (Instance& I : s.insts) {
        if (!I.d_fields2) continue;
        if (s.parity) std::swap(I.d_fields, I.d_fields2);
        (void)hipFree(I.d_fields2);
        I.d_fields2 = nullptr;
    }
    s.parity = 0;
    s.dirty = true;
    s.graph_dirty = true;
}

// A deterministic pseudo-random pattern in [-1e-3, 1e-3) over n elements: the autotuner's trials run
// on non-zero fields, whose stores cost what a real run's do (zero data has been measured faster).
template <typename T>
__global__ void __launch_bounds__(256) k_fill_pattern(T* p, int64_t n, uint32_t seed) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)i * 0x9e3779b1u ^ seed;
        h ^= h >> 16;
        h *= 0x85ebca6bu;
        h ^= h >> 13;
        p[i] = (T)((double)(h & 0xffffu) / 32768.0 - 1.0) * (T)1e-3;
    }
}

size_t monitor_bytes(const Session& s, const MonitorHost& h);
// bytes of one real component of a running-DFT accumulator or phase entry
inline size_t dft_element_size(const Session& s, const MonitorHost& h) {
    return (h.desc.flags & RFDTD_MONITOR_DFT64) ? sizeof(double) : s.esz;
}

// Launch autotuner: times every (variant, block size, x planes per workgroup) candidate over
// trial_steps field-only steps of this session's own grids and keeps the fastest.  The session must
// not have run yet (step 0); the trials run on a pseudo-random field pattern and the fields are zero
// afterwards.
template <typename T>
void autotune(Session& s, int trial_steps, int32_t cfg[3], double* ms_out) {
    if (s.step != 0) RF_THROW(RFDTD_ESTATE, "autotune a session before its first run (or after rfdtd_reset)");
    if (trial_steps < 1 || trial_steps > 10000) RF_THROW(RFDTD_EINVAL, "trial_steps must be in [1, 10000]");
    int maxNx = 1;
    for (const auto& M : s.members) maxNx = std::max(maxNx, M.Nx);
    std::vector<int> xps{0};  // 0: automatic (enough workgroups to fill the chip)
    for (int chunks : {2, 4, 8, 16, 32, 64, 128}) {
        const int xp = (maxNx + chunks - 1) / chunks;
        if (std::find(xps.begin(), xps.end(), xp) == xps.end()) xps.push_back(xp);
    }
    // one or two planes per workgroup: no marching, every plane's loads of a thread in flight at once
    // (the separate kernels on large grids run fastest so)
    for (int xp : {1, 2})
        if (std::find(xps.begin(), xps.end(), xp) == xps.end()) xps.push_back(xp);
    // the tiled passes also try chunk counts between the powers of two: how their tiles fill the
    // XCDs' rounds changes their time by several per cent (a trial each, not one per block size)
    std::vector<int> xps_tiled = xps;
    for (int chunks : {3, 5, 6, 10, 12}) {
        const int xp = (maxNx + chunks - 1) / chunks;
        if (std::find(xps_tiled.begin(), xps_tiled.end(), xp) == xps_tiled.end()) xps_tiled.push_back(xp);
    }
    const int v0 = variant_of(s);
    const int blk0 = s.block, ord0 = s.col_order;
    std::vector<int> xr0;
    for (const auto& M : s.members) xr0.push_back(M.xplanes_req);
    auto apply = [&](int v, int blk, const std::vector<int>& xr, int ord) {
        apply_variant(s, v);
        s.block = blk;
        s.col_order = ord;
        for (size_t m = 0; m < s.members.size(); ++m) s.members[m].xplanes_req = xr[m];
        s.dirty = true;
        s.graph_dirty = true;
    };
    hipStream_t st = s.own_stream;
    double best = 1e300, best_sep = 1e300;
    int bv = v0, bblk = blk0, bxp = 0, bord = ord0, sblk = s.sep_block, sxp = s.sep_xplanes, sord = s.sep_order;
    // With sources installed (and room for the trial steps) the trials apply them as a run does, so
    // that their cost is timed too; the probe rows and monitor values are restored afterwards and
    // the design-box history slots are rewritten by the run.  Otherwise field updates only.
    bool with_io = false, room = true;
    for (const Instance& I : s.insts) {
        with_io |= I.P > 0 || !I.mons.empty();
        if (I.P > 0 && s.Nt_probes < trial_steps) room = false;
        if (!I.mons.empty() && s.Nt_monitors < trial_steps) room = false;
    }
    for (const auto& h : s.hist) room &= h.capacity >= trial_steps;
    with_io &= room;
    s.tuning = !with_io;
    for (Instance& I : s.insts)
        k_fill_pattern<T><<<1024, 256, 0, st>>>((T*)I.d_fields, (int64_t)(I.fields_bytes / sizeof(T)), 12345u);
    HIP_OK(hipGetLastError());
    try {
        const bool separate_only = (s.flags & RFDTD_TUNE_SEPARATE) != 0;
        // every candidate with the workgroups in XCD ranges, and the separate kernels' streaming
        // shapes (one or two x planes per workgroup) in launch order; then the four fastest other
        // separate shapes again in launch order (it is faster on large grids, where they can win)
        std::vector<std::array<double, 4>> sep_times;
        struct Cand { int v, blk, xp, ord; };
        std::vector<Cand> cands;
        for (int v : {1, 4, 10, 11, 12, 0, 2, 3, 5, 6, 7, 8, 9, 13, 14, 15}) {
            const bool sep_v = !(v == 2 || v == 3 || v == 5 || v == 6 || is_two_step(v));
            for (int blk : {128, 256, 512})
                for (int xp : sep_v ? xps : xps_tiled) {
                    cands.push_back(Cand{v, blk, xp, 0});
                    if (sep_v && (xp == 1 || xp == 2)) cands.push_back(Cand{v, blk, xp, 1});
                }
        }
        const size_t n_main = cands.size();
        for (size_t ci = 0;; ++ci) {
            if (ci == n_main) {
                std::sort(sep_times.begin(), sep_times.end());
                int added = 0;
                for (size_t k = 0; k < sep_times.size() && added < 4; ++k) {
                    const int xp = (int)sep_times[k][3];
                    if (xp == 1 || xp == 2) continue;  // timed in launch order already
                    cands.push_back(Cand{(int)sep_times[k][1], (int)sep_times[k][2], xp, 1});
                    ++added;
                }
            }
            if (ci >= cands.size()) break;
            const int v = cands[ci].v, blk = cands[ci].blk, xp = cands[ci].xp, ord = cands[ci].ord;
            const bool fused_v = v == 2 || v == 3 || v == 5 || v == 6 || is_two_step(v);
            if (fused_v && separate_only) continue;
            if (!vec_fits(s, variant_vec(v))) continue;
            // a design-box history needs H after every step: the two-step pass would run
            // single steps
            if (is_two_step(v) && !s.hist.empty()) continue;
            if (fused_v && blk != 256) continue;  // the fused step's workgroup size is fixed by its tile
            apply(v, blk, std::vector<int>(s.members.size(), xp), ord);
            refresh<T>(s);
            run_chunk<T>(s, std::min(trial_steps, 3), 0, st, false, false, 0);  // warm-up
            HIP_OK(hipEventRecord(s.ev0, st));
            run_chunk<T>(s, trial_steps, 0, st, false, false, 0);
            HIP_OK(hipEventRecord(s.ev1, st));
            HIP_OK(hipEventSynchronize(s.ev1));
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
            const double per = (double)ms / trial_steps;
            if (s.flags & RFDTD_VERBOSE) {
                std::fprintf(stderr, "[rfdtd] autotune: variant %d, %d threads, x planes %d, %s: %.4f ms/step\n",
                             v, blk, xp, ord ? "launch order" : "XCD ranges", per);
                std::fflush(stderr);
            }
            if (per < best) {
                best = per;
                bv = v;
                bblk = blk;
                bxp = xp;
                bord = ord;
            }
            if (!fused_v && variant_vec(v) == 1 && per < best_sep) {  // the recomputation's shape
                best_sep = per;
                sblk = blk;
                sxp = xp;
                sord = ord;
            }
            if (!fused_v && ord == 0) sep_times.push_back({per, (double)v, (double)blk, (double)xp});
        }
    } catch (...) {
        s.tuning = false;
        apply(v0, blk0, xr0, ord0);
        throw;
    }
    s.tuning = false;
    apply(bv, bblk, std::vector<int>(s.members.size(), bxp), bord);
    s.sep_block = sblk;
    s.sep_xplanes = sxp;
    s.sep_order = sord;
    HIP_OK(hipDeviceSynchronize());
    for (Instance& I : s.insts) {
        HIP_OK(hipMemset(I.d_fields, 0, I.fields_bytes));
        if (I.d_fields2) HIP_OK(hipMemset(I.d_fields2, 0, I.fields_bytes));
        if (with_io) {
            if (I.P)
                HIP_OK(hipMemcpy(I.d_series, I.d_series_in, (size_t)s.Nt_probes * I.P * s.esz, hipMemcpyDeviceToDevice));
            for (auto& h : I.mons) HIP_OK(hipMemset(h.d_values, 0, monitor_bytes(s, h)));
        }
    }
    s.parity = 0;
    s.totals_stale = false;
    if (!s.fused) release_second_buffers(s);
    cfg[0] = bv;
    cfg[1] = bblk;
    cfg[2] = bxp;
    if (ms_out) *ms_out = best;
}

struct MemberSpec {
    int Nx, Ny, Nz, K;
    const void* const* coeff;
    const rfdtd_assembly* assembly = nullptr;  // coefficients assembled on the device instead
};

Session* make_session(const std::vector<MemberSpec>& specs, uint32_t flags) {
    if (specs.empty()) RF_THROW(RFDTD_EINVAL, "a session needs at least one grid");
    auto s = std::make_unique<Session>();
    s->f64 = (flags & RFDTD_F64) != 0;
    s->esz = s->f64 ? 8 : 4;
    s->flags = flags;
    HIP_OK(hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&s->ev0));
    HIP_OK(hipEventCreate(&s->ev1));
    HIP_OK(hipMalloc(&s->d_step_base, 256));
    int inst0 = 0;
    for (const MemberSpec& sp : specs) {
        if (sp.Nx < 2 || sp.Ny < 2 || sp.Nz < 2) RF_THROW(RFDTD_EINVAL, "grid must have at least 2 cells on every axis");
        // 32-bit element offsets in the kernels; byte offsets of double arrays must fit too
        if ((int64_t)sp.Nx * sp.Ny * sp.Nz > ((int64_t)1 << 29)) RF_THROW(RFDTD_EINVAL, "grid too large (max 2^29 cells)");
        if (sp.K < 1 || sp.K > 64) RF_THROW(RFDTD_EINVAL, "excitations per grid must be in [1, 64]");
        if (!sp.coeff && !sp.assembly) RF_THROW(RFDTD_EINVAL, "coefficients missing");
        Member M;
        M.Nx = sp.Nx; M.Ny = sp.Ny; M.Nz = sp.Nz;
        M.cells = (int64_t)sp.Nx * sp.Ny * sp.Nz;
        M.K = sp.K;
        M.inst0 = inst0;
        s->members.push_back(M);
        Member& Mr = s->members.back();
        if (sp.assembly) {
            if (s->f64) assemble_coefficients<double>(Mr, flags, *sp.assembly);
            else assemble_coefficients<float>(Mr, flags, *sp.assembly);
        } else if (s->f64) {
            upload_coefficients<double>(Mr, flags, sp.coeff);
        } else {
            upload_coefficients<float>(Mr, flags, sp.coeff);
        }
        for (int e = 0; e < sp.K; ++e) {
            Instance I;
            I.member = (int)s->members.size() - 1;
            I.pitch = align256((size_t)M.cells * s->esz);
            I.fields_bytes = 18 * I.pitch;
            HIP_OK(hipMalloc(&I.d_fields, I.fields_bytes));
            HIP_OK(hipMemset(I.d_fields, 0, I.fields_bytes));
            s->insts.push_back(std::move(I));
        }
        inst0 += sp.K;
    }
    // The launch shape a session starts with, before any set_variant or tuning: float sessions on the
    // coefficient table with at least kTwoStepDefaultWork cell-updates per step take the two-step
    // pass (26-row tiles, automatic x chunks): 1.78x faster than the separate kernels on the
    // 400x400x200 grid, 1.53x on 12 excitations of four port grids, 1.11x on 200x160x80 and on
    // 150x120x100; below that the separate kernels stay (as fast on 128x96x96, faster on smaller grids).
    int64_t work = 0;
    bool table = true;
    for (const Member& M : s->members) {
        work += M.cells * M.K;
        table &= M.coef_kind == 1;
    }
    if (!s->f64 && table && work >= kTwoStepDefaultWork) apply_variant(*s, 15);
    return s.release();
}

Instance& instance_at(Session& s, int k) {
    if (k < 0 || k >= (int)s.insts.size()) RF_THROW(RFDTD_EINVAL, "instance index out of range");
    return s.insts[k];
}

void set_probes(Session& s, int inst, const rfdtd_probe* probes, int n, int Nt) {
    if (n < 0 || n > RFDTD_MAX_PROBES) RF_THROW(RFDTD_EINVAL, "Exceeded maximum number of probes.");
    if (n > 0 && (!probes || Nt < 1)) RF_THROW(RFDTD_EINVAL, "invalid probe arguments");
    Instance& I = instance_at(s, inst);
    const Member& M = s.members[I.member];
    (void)hipDeviceSynchronize();
    s.release_probes(I);
    s.dirty = true;
    if (n == 0) return;
    bool other = false;
    for (const auto& J : s.insts) other |= (&J != &I) && J.P > 0;
    if (other && Nt != s.Nt_probes) RF_THROW(RFDTD_EINVAL, "all probe series of a session must have the same Nt");
    // component shifts of solver.cpp:503-520
    std::vector<int64_t> cell(n);
    std::vector<int32_t> field(n), is_src(n);
    for (int p = 0; p < n; ++p) {
        const rfdtd_probe& q = probes[p];
        if (q.field < 0 || q.field > 5) RF_THROW(RFDTD_EINVAL, "probe field must be 0..5");
        if (!q.values) RF_THROW(RFDTD_EINVAL, "probe values missing");
        int x = q.idx[0], y = q.idx[1], z = q.idx[2];
        const int f = q.field;
        if (f == EY || f == EZ || f == HX) x -= 1;
        if (f == EX || f == EZ || f == HY) y -= 1;
        if (f == EX || f == EY || f == HZ) z -= 1;
        if (x < 0 || y < 0 || z < 0) RF_THROW(RFDTD_EINVAL, "Source cannot be placed on grid edge");
        if (x >= M.Nx || y >= M.Ny || z >= M.Nz) RF_THROW(RFDTD_EINVAL, "probe index outside the grid");
        cell[p] = ((int64_t)x * M.Ny + y) * M.Nz + z;
        field[p] = f;
        is_src[p] = q.is_source ? 1 : 0;
    }
    // groups by (component, cell) in order of first appearance; E groups first
    std::vector<int32_t> order;
    for (int half = 0; half < 2; ++half) {
        std::unordered_map<int64_t, int> gid;
        std::vector<std::vector<int32_t>> members;
        std::vector<Instance::HostGroup> gs;
        for (int p = 0; p < n; ++p) {
            if ((field[p] >= 3) != (half == 1)) continue;
            const int64_t key = cell[p] * 8 + field[p];
            auto it = gid.find(key);
            if (it == gid.end()) {
                it = gid.emplace(key, (int)gs.size()).first;
                gs.push_back(Instance::HostGroup{cell[p], field[p], 0, 0});
                members.emplace_back();
            }
            members[it->second].push_back(p);
        }
        for (size_t g = 0; g < gs.size(); ++g) {
            gs[g].start = (int32_t)order.size();
            gs[g].count = (int32_t)members[g].size();
            order.insert(order.end(), members[g].begin(), members[g].end());
        }
        if (half == 0) I.ngE = (int)gs.size();
        I.groups.insert(I.groups.end(), gs.begin(), gs.end());
    }
    I.P = n;
    s.Nt_probes = Nt;
    // time-major series [Nt][P]
    std::vector<char> series((size_t)Nt * n * s.esz);
    for (int p = 0; p < n; ++p) {
        const char* src = (const char*)probes[p].values;
        for (int t = 0; t < Nt; ++t)
            std::memcpy(&series[((size_t)t * n + p) * s.esz], src + (size_t)t * s.esz, s.esz);
        I.probe_host_values.push_back(probes[p].values);
    }
    HIP_OK(hipMalloc(&I.d_series, series.size()));
    HIP_OK(hipMemcpy(I.d_series, series.data(), series.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&I.d_series_in, series.size()));
    HIP_OK(hipMemcpy(I.d_series_in, series.data(), series.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&I.d_order, order.size() * 4));
    HIP_OK(hipMemcpy(I.d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&I.d_is_src, (size_t)n * 4));
    HIP_OK(hipMemcpy(I.d_is_src, is_src.data(), (size_t)n * 4, hipMemcpyHostToDevice));
}

void set_monitors(Session& s, int inst, const rfdtd_monitor* mons, int n, int Nt) {
    if (n < 0 || n > RFDTD_MAX_MONITORS) RF_THROW(RFDTD_EINVAL, "Exceeded maximum number of monitors.");
    if (n > 0 && (!mons || Nt < 1)) RF_THROW(RFDTD_EINVAL, "invalid monitor arguments");
    Instance& I = instance_at(s, inst);
    const Member& M = s.members[I.member];
    (void)hipDeviceSynchronize();
    s.release_monitors(I);
    s.dirty = true;
    if (n == 0) return;
    bool other = false;
    for (const auto& J : s.insts) other |= (&J != &I) && !J.mons.empty();
    if (other && Nt != s.Nt_monitors) RF_THROW(RFDTD_EINVAL, "all monitors of a session must have the same Nt");
    s.Nt_monitors = Nt;
    for (int m = 0; m < n; ++m) {
        const rfdtd_monitor& d = mons[m];
        if (d.axis < 0 || d.axis > 2 || d.field < 0 || d.field > 5 || d.n_step < 1 || d.n_frequencies < 0)
            RF_THROW(RFDTD_EINVAL, "invalid monitor description");
        if (!d.values || (d.n_frequencies > 0 && !d.dtft_phase)) RF_THROW(RFDTD_EINVAL, "monitor arrays missing");
        if (d.flags & ~(int32_t)RFDTD_MONITOR_DFT64) RF_THROW(RFDTD_EINVAL, "unknown monitor flags");
        if ((d.flags & RFDTD_MONITOR_DFT64) && d.n_frequencies == 0)
            RF_THROW(RFDTD_EINVAL, "RFDTD_MONITOR_DFT64 is for running-DFT monitors");
        MonitorHost h;
        h.desc = d;
        const int f = d.field;
        int pos = d.position;
        // GPU-packing shifts of solver.cpp:394-433 (gpu = 1)
        if (d.axis == 0 && (f == HX || f == EY || f == EZ)) pos -= 1;
        if (d.axis == 1 && (f == HY || f == EX || f == EZ)) pos -= 1;
        if (d.axis == 2 && (f == HZ || f == EX || f == EY)) pos -= 1;
        h.pos = pos;
        const int along = d.axis == 0 ? M.Nx : d.axis == 1 ? M.Ny : M.Nz;
        h.N1 = d.axis == 0 ? M.Ny : M.Nx;
        h.N2 = d.axis == 2 ? M.Ny : M.Nz;
        h.Nm = Nt / d.n_step + 1;
        h.active = pos >= 0 && pos < along;
        if (s.stream_ring > 0 && !d.n_frequencies && h.active) h.ring = std::min(h.Nm, s.stream_ring);
        const size_t vbytes = monitor_bytes(s, h);
        HIP_OK(hipMalloc(&h.d_values, vbytes));
        if (h.ring) HIP_OK(hipMemset(h.d_values, 0, vbytes));  // the host array keeps unsampled rows
        else HIP_OK(hipMemcpy(h.d_values, d.values, vbytes, hipMemcpyHostToDevice));
        if (d.n_frequencies) {
            const size_t pbytes = (size_t)h.Nm * d.n_frequencies * 2 * dft_element_size(s, h);
            HIP_OK(hipMalloc(&h.d_phase, pbytes));
            HIP_OK(hipMemcpy(h.d_phase, d.dtft_phase, pbytes, hipMemcpyHostToDevice));
        }
        I.mons.push_back(h);
    }
}

void read_probes(Session& s) {
    HIP_OK(hipDeviceSynchronize());
    for (Instance& I : s.insts) {
        if (!I.P) continue;
        std::vector<char> series((size_t)s.Nt_probes * I.P * s.esz);
        HIP_OK(hipMemcpy(series.data(), I.d_series, series.size(), hipMemcpyDeviceToHost));
        for (int p = 0; p < I.P; ++p) {
            char* dst = (char*)I.probe_host_values[p];
            for (int t = 0; t < s.Nt_probes; ++t)
                std::memcpy(dst + (size_t)t * s.esz, &series[((size_t)t * I.P + p) * s.esz], s.esz);
        }
    }
}

size_t monitor_bytes(const Session& s, const MonitorHost& h) {
    const size_t plane = (size_t)h.N1 * h.N2;
    return h.desc.n_frequencies ? (size_t)h.desc.n_frequencies * plane * 2 * dft_element_size(s, h)
                                : (size_t)(h.ring ? h.ring : h.Nm) * plane * s.esz;
}

void read_monitors(Session& s) {
    HIP_OK(hipDeviceSynchronize());
    if (s.copy_stream) HIP_OK(hipStreamSynchronize(s.copy_stream));
    for (Instance& I : s.insts)
        for (auto& h : I.mons)
            if (!h.ring) HIP_OK(hipMemcpy(h.desc.values, h.d_values, monitor_bytes(s, h), hipMemcpyDeviceToHost));
}

void read_fields(Session& s, int inst, void* const out[18]) {
    HIP_OK(hipDeviceSynchronize());
    Instance& I = instance_at(s, inst);
    const Member& M = s.members[I.member];
    const int Nx = M.Nx, Ny = M.Ny, Nz = M.Nz;
    std::vector<char> buf((size_t)M.cells * s.esz), b2;
    // component of each of the 18 arrays and the offset of GPU cell (0,0,0) in the global array
    const int comp[18] = {EX, EX, EY, EY, EZ, EZ, HX, HX, HY, HY, HZ, HZ, EX, EY, EZ, HX, HY, HZ};
    const char* cur = (const char*)s.cur(I);
    for (int a = 0; a < 18; ++a) {
        if (!out[a]) continue;
        if (a < 12) {
            HIP_OK(hipMemcpy(buf.data(), cur + (size_t)a * I.pitch, buf.size(), hipMemcpyDeviceToHost));
        } else {
            // a total is the sum of its two splits, formed as the update kernels form it
            const int k = 2 * (a - 12);
            b2.resize(buf.size());
            HIP_OK(hipMemcpy(buf.data(), cur + (size_t)k * I.pitch, buf.size(), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(b2.data(), cur + (size_t)(k + 1) * I.pitch, b2.size(), hipMemcpyDeviceToHost));
            if (s.f64)
                for (int64_t i = 0; i < M.cells; ++i) ((double*)buf.data())[i] += ((const double*)b2.data())[i];
            else
                for (int64_t i = 0; i < M.cells; ++i) ((float*)buf.data())[i] += ((const float*)b2.data())[i];
        }
        const int c = comp[a];
        const int gx = Nx + ((c == EY || c == EZ || c == HX) ? 1 : 0);
        const int gy = Ny + ((c == EX || c == EZ || c == HY) ? 1 : 0);
        const int gz = Nz + ((c == EX || c == EY || c == HZ) ? 1 : 0);
        const int ox = gx - Nx, oy = gy - Ny, oz = gz - Nz;
        char* dst = (char*)out[a];
        std::memset(dst, 0, (size_t)gx * gy * gz * s.esz);
        for (int x = 0; x < Nx; ++x)
            for (int y = 0; y < Ny; ++y)
                std::memcpy(dst + (((size_t)(x + ox) * gy + (y + oy)) * gz + oz) * s.esz,
                            &buf[(((size_t)x * Ny + y) * Nz) * s.esz], (size_t)Nz * s.esz);
    }
}

// Device buffers of one gradient evaluation, freed on every exit path.
struct GradBuffers {
    std::vector<void*> ptrs;
    void* alloc(size_t bytes) {
        void* p = nullptr;
        HIP_OK(hipMalloc(&p, bytes));
        ptrs.push_back(p);
        return p;
    }
    ~GradBuffers() {
        (void)hipDeviceSynchronize();
        for (void* p : ptrs) (void)hipFree(p);
    }
};


// One gradient request: instance, design box, dJ/dy of its probes, output.
struct GradItem {
    int inst;
    Box B;
    const void* const* adj;
    int n_adj;
    double* out;
};

// Reverse-time pass over a checkpointed forward run, for several instances in the same launches:
// each checkpoint segment is recomputed forward from its checkpoint (saving H in every design box
// each step), then the adjoint steps back through the segment and accumulates dJ/dCb of the six E
// splits in every box.  adj[p] of an item is dJ/dy of probe p, T[Nt] in the instance's probe
// order; out: double[6][nx][ny][nz] (ex_y ex_z ey_z ey_x ez_x ez_y), box cells in grid coordinates.
template <typename T>
void gradients(Session& s, const std::vector<GradItem>& items) {
    const int64_t Nt = s.step;
    if (Nt < 1) RF_THROW(RFDTD_ESTATE, "run the forward solve first");
    const int n = (int)items.size();
    if (n < 1) RF_THROW(RFDTD_EINVAL, "no gradient requested");
    // history mode: the forward run recorded every item's box H at every step, so the reverse pass
    // needs no checkpoints and no recomputation
    std::vector<const Session::History*> hs(n, nullptr);
    bool hist_mode = true;
    for (int k = 0; k < n; ++k) {
        const GradItem& it = items[k];
        for (const auto& h : s.hist)
            if (h.inst == it.inst && h.capacity >= Nt && h.box[0] == it.B.x0 && h.box[2] == it.B.y0 &&
                h.box[4] == it.B.z0 && h.box[1] - h.box[0] == it.B.nx && h.box[3] - h.box[2] == it.B.ny &&
                h.box[5] - h.box[4] == it.B.nz)
                hs[k] = &h;
        hist_mode &= hs[k] != nullptr;
    }
    if (!hist_mode && s.ckpt_every <= 0)
        RF_THROW(RFDTD_ESTATE, "record the design-box history or enable checkpointing before the forward run");
    const int K = hist_mode ? (int)std::min<int64_t>(Nt, INT32_MAX) : s.ckpt_every;
    const int nseg = hist_mode ? 1 : (int)((Nt + K - 1) / K);
    std::vector<int> seen(s.insts.size(), 0);
    for (const GradItem& it : items) {
        Instance& I = instance_at(s, it.inst);
        const Member& M = s.members[I.member];
        if (seen[it.inst]++) RF_THROW(RFDTD_EINVAL, "an instance may appear once per gradient call");
        if (!hist_mode && (int)I.ckpt.size() < nseg) RF_THROW(RFDTD_ESTATE, "missing checkpoints of the forward run");
        if (it.n_adj != I.P || (I.P && !it.adj)) RF_THROW(RFDTD_EINVAL, "one adjoint series per probe is required");
        const Box& B = it.B;
        if (B.nx < 1 || B.ny < 1 || B.nz < 1 || B.x0 < 0 || B.y0 < 0 || B.z0 < 0 || B.x0 + B.nx > M.Nx ||
            B.y0 + B.ny > M.Ny || B.z0 + B.nz > M.Nz)
            RF_THROW(RFDTD_EINVAL, "design box outside the grid");
        if (!it.out) RF_THROW(RFDTD_EINVAL, "gradient output missing");
    }
    refresh<T>(s);
    HIP_OK(hipDeviceSynchronize());
    hipStream_t st = s.own_stream;
    HIP_OK(hipEventRecord(s.ev0, st));
    const size_t esz = sizeof(T);
    GradBuffers buf;
    // descriptors: [0, n) recomputed forward fields, [n, 2n) adjoint fields
    std::vector<InstDev<T>> di(2 * n);
    std::vector<MemberDev> md(n);
    std::vector<void*> d_re(n);
    std::vector<BoxDev<T>> bd(n);
    std::vector<int64_t> cells(n), hcells(n), gcells(n);
    std::vector<ProbeGroup<T>> gE, gH, aE, aH;
    // the recomputation runs the separate kernels at their own launch shape (sep_block threads; x
    // planes as tuned, else enough workgroups to fill the chip), whatever variant ran the forward
    const int rb = s.sep_block;
    int64_t rbpp = 0;
    for (const GradItem& it : items) {
        const Member& M = s.members[s.insts[it.inst].member];
        rbpp += ((int64_t)M.Ny * M.Nz + rb - 1) / rb;
    }
    auto rxplanes = [&](const Member& M) {
        if (s.sep_xplanes > 0) return std::min(s.sep_xplanes, M.Nx);
        const int64_t chunks = std::min<int64_t>(M.Nx, std::max<int64_t>(1, (2048 + rbpp - 1) / rbpp));
        return (int)((M.Nx + chunks - 1) / chunks);
    };
    int tile0 = 0;
    for (int k = 0; k < n; ++k) {
        const GradItem& it = items[k];
        Instance& I = s.insts[it.inst];
        const Member& M = s.members[I.member];
        void* d_adj = buf.alloc(18 * I.pitch);
        d_re[k] = hist_mode ? d_adj : buf.alloc(18 * I.pitch);  // no recomputation in history mode
        HIP_OK(hipMemsetAsync(d_adj, 0, 18 * I.pitch, st));
        const int npb = (int)(((int64_t)M.Ny * M.Nz + rb - 1) / rb);
        const int rxp = rxplanes(M);
        const int chunks = (M.Nx + rxp - 1) / rxp;
        di[k] = make_instdev<T>(M, d_re[k], I.pitch, npb);
        di[k].xplanes = rxp;
        // the recomputation applies the sources inside the update kernels, from the pristine series;
        // the records it writes equal the forward run's (the same operations on the same state)
        di[k].ebits = I.d_ebits;
        di[k].egroups = I.d_egroups;
        di[k].negroups = I.negroups;
        di[k].hbits = I.d_hbits;
        di[k].hgroups = I.d_hgroups;
        di[k].nhgroups = I.nhgroups;
        di[n + k] = make_instdev<T>(M, d_adj, I.pitch, npb);
        md[k] = MemberDev{tile0, npb * chunks, 1, k, s.sep_order};
        tile0 += npb * chunks;
        cells[k] = M.cells;
        // adjoint sources, time-major [Nt][P]
        T* d_g = nullptr;
        if (I.P) {
            std::vector<T> gser((size_t)Nt * I.P);
            for (int p = 0; p < I.P; ++p) {
                if (!it.adj[p]) RF_THROW(RFDTD_EINVAL, "adjoint series missing");
                const T* src = (const T*)it.adj[p];
                for (int64_t t = 0; t < Nt; ++t) gser[(size_t)t * I.P + p] = src[t];
            }
            d_g = (T*)buf.alloc(gser.size() * esz);
            HIP_OK(hipMemcpyAsync(d_g, gser.data(), gser.size() * esz, hipMemcpyHostToDevice, st));
            // the copy reads host memory that goes out of scope: finish it now
            HIP_OK(hipStreamSynchronize(st));
        }
        for (size_t g = 0; g < I.groups.size(); ++g) {
            const auto& hg = I.groups[g];
            ProbeGroup<T> pg{};
            pg.cell = hg.cell;
            pg.series = (T*)I.d_series;
            pg.series_in = (const T*)I.d_series_in;
            pg.order = I.d_order + hg.start;
            pg.is_src = I.d_is_src;
            pg.inst = k;
            pg.field = hg.field;
            pg.count = hg.count;
            pg.nprobes = I.P;
            ProbeGroup<T> ag = pg;
            ag.series = d_g;
            ((int)g < I.ngE ? gE : gH).push_back(pg);
            ((int)g < I.ngE ? aE : aH).push_back(ag);
        }
        const Box& B = it.B;
        const int64_t slot_elems = (int64_t)3 * (B.nx + 1) * (B.ny + 1) * (B.nz + 1);
        bd[k].B = B;
        bd[k].inst = k;
        bd[k].slot_elems = slot_elems;
        bd[k].slots = hist_mode ? (T*)hs[k]->d : (T*)buf.alloc((size_t)K * slot_elems * esz);
        const size_t nbox = (size_t)B.nx * B.ny * B.nz;
        bd[k].grad = (double*)buf.alloc(6 * nbox * sizeof(double));
        HIP_OK(hipMemsetAsync(bd[k].grad, 0, 6 * nbox * sizeof(double), st));
        hcells[k] = slot_elems / 3;
        gcells[k] = (int64_t)nbox;
    }
    const std::vector<int64_t> flat0 = block_offsets(cells), h0 = block_offsets(hcells);
    auto upload = [&](const void* src, size_t bytes) {
        void* d = buf.alloc(bytes);
        HIP_OK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, st));
        return d;
    };
    // reverse-pass tiling: its own workgroup size and planes per workgroup in the adjoint
    // descriptors [n, 2n) and member table d_mda
    std::vector<MemberDev> mda(n);
    auto adj_tiling = [&](int block, int chunks, int order) {
        int tiles = 0;
        for (int k = 0; k < n; ++k) {
            const Member& M = s.members[s.insts[items[k].inst].member];
            const int xp = chunks > 0 ? (M.Nx + chunks - 1) / chunks : rxplanes(M);
            const int npb = (int)(((int64_t)M.Ny * M.Nz + block - 1) / block);
            const int nch = (M.Nx + xp - 1) / xp;
            di[n + k].npb = npb;
            di[n + k].xplanes = xp;
            mda[k] = MemberDev{tiles, npb * nch, 1, k, order};
            tiles += npb * nch;
        }
        return tiles;
    };
    // each item's adjoint sources sorted by cell, with a bit per cell, for the reverse kernels
    for (int k = 0; k < n; ++k) {
        const Member& M = s.members[s.insts[items[k].inst].member];
        for (int h = 0; h < 2; ++h) {
            std::vector<ProbeGroup<T>> g;
            for (const auto& pg : h ? aH : aE)
                if (pg.inst == k) g.push_back(pg);
            if (g.empty()) continue;
            std::stable_sort(g.begin(), g.end(),
                             [](const ProbeGroup<T>& a, const ProbeGroup<T>& b) { return a.cell < b.cell; });
            std::vector<uint32_t> bits((size_t)(M.cells + 31) / 32, 0u);
            for (const auto& pg : g) bits[(size_t)(pg.cell >> 5)] |= 1u << (pg.cell & 31);
            const uint32_t* d_bits = (const uint32_t*)upload(bits.data(), bits.size() * 4);
            const void* d_grp = upload(g.data(), g.size() * sizeof(ProbeGroup<T>));
            HIP_OK(hipStreamSynchronize(st));  // the copies read bits and g
            InstDev<T>& d = di[n + k];
            if (h) {
                d.hbits = d_bits; d.hgroups = d_grp; d.nhgroups = (int)g.size();
            } else {
                d.ebits = d_bits; d.egroups = d_grp; d.negroups = (int)g.size();
            }
        }
    }
    int ablock = s.adj_block > 0 ? s.adj_block : rb;
    // the reverse pass's variant: bit 0 pipelined columns, bit 1 workgroups in launch order
    int atiles = adj_tiling(ablock, s.adj_block > 0 ? s.adj_chunks : 0, s.adj_pf >> 1);
    const InstDev<T>* d_di = (const InstDev<T>*)upload(di.data(), di.size() * sizeof(InstDev<T>));
    const MemberDev* d_md = (const MemberDev*)upload(md.data(), md.size() * sizeof(MemberDev));
    MemberDev* d_mda = (MemberDev*)upload(mda.data(), mda.size() * sizeof(MemberDev));
    // step >= 0: the kernel also adds that step's adjoint sources (E for adj_e, H for adj_h)
    auto adj_e = [&](int pf, int block, int tiles, int step) {
        const int g = (tiles + 7) & ~7;
        if (pf) k_adj_e<T, true><<<g, block, 0, st>>>(d_di + n, d_mda, n, tiles, step);
        else k_adj_e<T, false><<<g, block, 0, st>>>(d_di + n, d_mda, n, tiles, step);
    };
    auto adj_h = [&](int pf, int block, int tiles, int step) {
        const int g = (tiles + 7) & ~7;
        if (pf) k_adj_h<T, true><<<g, block, 0, st>>>(d_di + n, d_mda, n, tiles, step);
        else k_adj_h<T, false><<<g, block, 0, st>>>(d_di + n, d_mda, n, tiles, step);
    };
    if (s.adj_block <= 0) {
        HIP_OK(hipStreamSynchronize(st));  // the descriptor uploads above read di and mda
        // time the candidate shapes on a pseudo-random adjoint state (zero data runs faster than
        // real data; the state is cleared again below) and keep the fastest for this session
        std::vector<int> chunk_opts;
        int maxNx = 1;
        for (const GradItem& it : items) maxNx = std::max(maxNx, s.members[s.insts[it.inst].member].Nx);
        for (int c : {1, 2, 4, 8, 16, 32, 64, 128, 256})
            if (c < maxNx) chunk_opts.push_back(c);
        chunk_opts.push_back(maxNx);  // one plane per workgroup
        for (int k = 0; k < n; ++k)
            k_fill_pattern<T><<<1024, 256, 0, st>>>((T*)di[n + k].F.ex_y,
                                                     (int64_t)(18 * s.insts[items[k].inst].pitch / sizeof(T)), 777u);
        double best = 1e300;
        int bpf = 1, bblk = ablock, bch = 1;
        const int reps = 3;
        for (int av : {1, 0, 3, 2})
            for (int blk : {128, 256, 512})
                for (int ch : chunk_opts) {
                    const int pf = av & 1, ord = av >> 1;
                    if (ord && 2 * ch < maxNx) continue;  // launch order pays on the streaming shapes
                    const int tiles = adj_tiling(blk, ch, ord);
                    HIP_OK(hipMemcpyAsync((void*)(d_di + n), di.data() + n, n * sizeof(InstDev<T>), hipMemcpyHostToDevice, st));
                    HIP_OK(hipMemcpyAsync(d_mda, mda.data(), n * sizeof(MemberDev), hipMemcpyHostToDevice, st));
                    adj_e(pf, blk, tiles, -1);
                    adj_h(pf, blk, tiles, -1);
                    HIP_OK(hipEventRecord(s.ev0, st));
                    for (int r = 0; r < reps; ++r) {
                        adj_e(pf, blk, tiles, -1);
                        adj_h(pf, blk, tiles, -1);
                    }
                    HIP_OK(hipEventRecord(s.ev1, st));
                    HIP_OK(hipEventSynchronize(s.ev1));
                    HIP_OK(hipGetLastError());
                    float ms = 0;
                    HIP_OK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
                    if (ms < best) {
                        best = ms;
                        bpf = av;
                        bblk = blk;
                        bch = ch;
                    }
                }
        s.adj_pf = bpf;
        s.adj_block = bblk;
        s.adj_chunks = bch;
        if (s.flags & RFDTD_VERBOSE) {
            std::fprintf(stderr, "[rfdtd] adjoint launch: %s columns, %s, %d threads, %d x chunks (%.3f ms per step)\n",
                         (bpf & 1) ? "pipelined" : "plain", (bpf >> 1) ? "launch order" : "XCD ranges", bblk, bch,
                         best / reps);
            std::fflush(stderr);
        }
        ablock = bblk;
        atiles = adj_tiling(ablock, bch, bpf >> 1);
        HIP_OK(hipMemcpyAsync((void*)(d_di + n), di.data() + n, n * sizeof(InstDev<T>), hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(d_mda, mda.data(), n * sizeof(MemberDev), hipMemcpyHostToDevice, st));
        for (int k = 0; k < n; ++k) HIP_OK(hipMemsetAsync(di[n + k].F.ex_y, 0, 18 * s.insts[items[k].inst].pitch, st));
        HIP_OK(hipEventRecord(s.ev0, st));
    }
    const int apf = s.adj_pf & 1;
    const BoxDev<T>* d_bd = (const BoxDev<T>*)upload(bd.data(), bd.size() * sizeof(BoxDev<T>));
    const int64_t* d_flat0 = (const int64_t*)upload(flat0.data(), flat0.size() * 8);
    const int64_t* d_h0 = (const int64_t*)upload(h0.data(), h0.size() * 8);
    // the gradient's workgroups in the H adjoint launch: offsets in workgroups of its size
    const std::vector<int64_t> ga0 = block_offsets(gcells, ablock);
    const int64_t* d_ga0 = (const int64_t*)upload(ga0.data(), ga0.size() * 8);
    const int gab = (int)ga0.back();
    const int ngE = (int)gE.size(), ngH = (int)gH.size();  // the adjoint groups' E / H split
    aE.insert(aE.end(), aH.begin(), aH.end());
    const ProbeGroup<T>* d_ag = aE.empty() ? nullptr : (const ProbeGroup<T>*)upload(aE.data(), aE.size() * sizeof(ProbeGroup<T>));
    HIP_OK(hipStreamSynchronize(st));  // host-side descriptor vectors stay alive; uploads done
    const int grid = (tile0 + 7) & ~7;
    const int flat = (int)flat0.back(), hblocks = (int)h0.back();
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int64_t n0 = hist_mode ? 0 : (int64_t)seg * K, n1 = hist_mode ? Nt : std::min<int64_t>(Nt, n0 + K);
        if (!hist_mode)
            for (int k = 0; k < n; ++k) {
                const Instance& I = s.insts[items[k].inst];
                HIP_OK(hipMemcpyAsync(d_re[k], I.ckpt[seg], 12 * I.pitch, hipMemcpyDeviceToDevice, st));
            }
        if (!hist_mode) k_totals<T><<<flat, 256, 0, st>>>(d_di, d_flat0, n);
        for (int64_t t = n0; t < n1 && !hist_mode; ++t) {
            const int j = (int)(t - n0);
            k_save_h<T><<<hblocks, 256, 0, st>>>(d_di, d_bd, d_h0, n, j);
            k_update_e<T, true><<<grid, rb, 0, st>>>(d_di, d_md, n, tile0, nullptr, (int)t, 1);
            k_update_h<T, true><<<grid, rb, 0, st>>>(d_di, d_md, n, tile0, nullptr, (int)t, 1);
        }
        for (int64_t t = n1 - 1; t >= n0; --t) {
            const int j = (int)(t - n0);
            // the H adjoint sources of step t were added by the previous reverse step's adj_h (the
            // last step's here); those of E by adj_e itself
            if (ngH && t == Nt - 1) k_adj_inject<T><<<(ngH + 63) / 64, 64, 0, st>>>(d_di + n, d_ag + ngE, ngH, (int)t);
            adj_e(apf, ablock, atiles, (int)t);
            // H adjoint and the box gradient of step t (which reads the E adjoint adj_e just wrote
            // and the saved H^t, neither written by the H pass) in one launch
            const int g = (atiles + gab + 7) & ~7;
            if (apf) k_adj_h<T, true><<<g, ablock, 0, st>>>(d_di + n, d_mda, n, atiles, (int)t - 1, d_bd, d_ga0, n, j, d_di + n);
            else k_adj_h<T, false><<<g, ablock, 0, st>>>(d_di + n, d_mda, n, atiles, (int)t - 1, d_bd, d_ga0, n, j, d_di + n);
        }
        HIP_OK(hipGetLastError());
        if ((s.flags & RFDTD_VERBOSE) && (seg % std::max(1, nseg / 10) == 0 || seg == 0)) {
            HIP_OK(hipStreamSynchronize(st));
            std::fprintf(stderr, "[rfdtd] adjoint: %d of %d checkpoint segments left\n", seg, nseg);
            std::fflush(stderr);
        }
    }
    HIP_OK(hipEventRecord(s.ev1, st));
    for (int k = 0; k < n; ++k)
        HIP_OK(hipMemcpyAsync(items[k].out, bd[k].grad, 6 * (size_t)gcells[k] * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    float ms = 0;
    if (hipEventElapsedTime(&ms, s.ev0, s.ev1) == hipSuccess) s.last_total_ms = ms;
    else (void)hipGetLastError();  // clear the error the failed query leaves
}

// Checkpoint interval for a gradient over `items` after an Nt-step forward run: the interval K that
// minimises checkpoint memory (ceil(Nt/K) sets of every instance's 12 split arrays) plus the design
// boxes' per-step H slots (K of them), checked against `budget` bytes (0: the device's free memory
// less 1 GiB).
template <typename T>
void plan_checkpoints(Session& s, int Nt, const std::vector<GradItem>& items, uint64_t budget, int32_t* every_out,
                      uint64_t* bytes_out) {
    if (Nt < 1) RF_THROW(RFDTD_EINVAL, "Nt must be >= 1");
    const size_t esz = sizeof(T);
    double per_ckpt = 0, per_slot = 0, fixed = 0;
    for (const Instance& I : s.insts) per_ckpt += 12.0 * I.pitch;
    for (const GradItem& it : items) {
        const Instance& I = instance_at(s, it.inst);
        const Box& B = it.B;
        if (B.nx < 1 || B.ny < 1 || B.nz < 1) RF_THROW(RFDTD_EINVAL, "empty design box");
        per_slot += 3.0 * (B.nx + 1) * (B.ny + 1) * (B.nz + 1) * esz;
        fixed += 36.0 * I.pitch + (double)Nt * I.P * esz + 48.0 * B.nx * B.ny * B.nz;
    }
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        budget = free_b > (1ull << 30) ? free_b - (1ull << 30) : 0;
    }
    int best = 1;
    double best_bytes = 1e300;
    for (int K = 1; K <= Nt; ++K) {
        const double b = std::ceil((double)Nt / K) * per_ckpt + (double)K * per_slot + fixed;
        if (b < best_bytes) {
            best_bytes = b;
            best = K;
        }
    }
    if (best_bytes > (double)budget)
        RF_THROW(RFDTD_ENOMEM, "checkpoints and adjoint buffers need " + std::to_string((uint64_t)best_bytes) +
                                   " bytes, more than the " + std::to_string(budget) + " available");
    *every_out = best;
    if (bytes_out) *bytes_out = (uint64_t)best_bytes;
}

void progress_bar(int64_t n, int64_t Nt) {
    const int pct = (int)((n + 1) * 100 / Nt);
    const int dots = pct / 5;
    std::string str = "\r|";
    str.append(dots, '.');
    str.append(20 - dots, ' ');
    str += "| " + std::to_string(pct) + "%  ";
    std::fputs(str.c_str(), stdout);
    std::fflush(stdout);
}

template <class Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        g_last_error.clear();
        return RFDTD_OK;
    } catch (RfdtdError& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return RFDTD_EINVAL;
    } catch (std::exception& e) {
        g_last_error = e.what();
        return RFDTD_EINVAL;
    }
}

Session& S_of(rfdtd_session* s) {
    if (!s) RF_THROW(RFDTD_EINVAL, "null session");
    return *(Session*)s;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// The far field without device allocation: the staged inputs and the result live in a file-scope
// device array, so that callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid it) can run the radiation integrals.  Every GPU the library
// runs on holds its own copy of the array until the process ends.  In host code `scratch` names a
// host-side stand-in; its device address on the current GPU comes from hipGetSymbolAddress.
// ------------------------------------------------------------------------------------------------

constexpr size_t kFFScratchBytes = (size_t)64 << 20;  // the workspace of rfdtd_nf2ff_static
constexpr int kFFMaxFaces = 64;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kFFScratchBytes];

namespace {

std::mutex g_scratch_mu;                // held for a whole rfdtd_nf2ff_static call: calls share the array
void* g_scratch_dev[kMaxDevices] = {};  // the array's device address on each GPU, taken once per device

// The workspace's device address on the current GPU (looked up on the device's first call, with the
// device current).  Called with g_scratch_mu held.
char* scratch_address() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) RF_THROW(RFDTD_EDEVICE, "device index out of range");
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) RF_THROW(RFDTD_EDEVICE, "no device address for the far-field workspace");
        g_scratch_dev[dev] = p;
    }
    return (char*)g_scratch_dev[dev];
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------

namespace {

using DesignHeld = Member::DesignHeld;

Member& design_member(Session& S, int member) {
    if (member < 0 || member >= (int)S.members.size()) RF_THROW(RFDTD_EINVAL, "member out of range");
    return S.members[member];
}

// Validates a design box of member M and the permittivity it holds, and puts the box's arrays on the
// device (one allocation, returned in .buf, which the caller owns): the tuples of eps_now must equal
// the session's coefficients, else RFDTD_EINVAL (a box touching PML, conductor, lumped-element or
// edge-corrected edges, or a permittivity that is not the session's).  In table form the box's cells
// are given entries of their own: the ones they already have if this box was set before, else new
// ones after every entry in use (cells of earlier boxes keep theirs).
DesignHeld design_hold(Session& S, Member& M, const int32_t* gbox, const double* eps_now, const double* sigma,
                       const double* dx, const double* dy, const double* dz, const void* dx1, const void* dy1,
                       const void* dz1, double dt, double e0) {
    if (!gbox || !eps_now || !sigma || !dx || !dy || !dz || !dx1 || !dy1 || !dz1)
        RF_THROW(RFDTD_EINVAL, "design permittivity arrays missing");
    const int gx0 = gbox[0], gx1 = gbox[1], gy0 = gbox[2], gy1 = gbox[3], gz0 = gbox[4], gz1 = gbox[5];
    if (gx0 < 0 || gy0 < 0 || gz0 < 0 || gx1 <= gx0 || gy1 <= gy0 || gz1 <= gz0 || gx1 > M.Nx || gy1 > M.Ny ||
        gz1 > M.Nz)
        RF_THROW(RFDTD_EINVAL, "design box outside the grid");
    const int64_t n = (int64_t)(gx1 - gx0) * (gy1 - gy0) * (gz1 - gz0);
    if (n > (int64_t)UINT32_MAX - M.nE - M.design_used) RF_THROW(RFDTD_EINVAL, "design box too large");
    DesignHeld H;
    DesignDev& D = H.D;
    D.gx0 = gx0; D.gy0 = gy0; D.gz0 = gz0;
    D.gnx = gx1 - gx0; D.gny = gy1 - gy0; D.gnz = gz1 - gz0;
    D.ex1 = std::min(gx1 + 1, M.Nx) - gx0;
    D.ey1 = std::min(gy1 + 1, M.Ny) - gy0;
    D.ez1 = std::min(gz1 + 1, M.Nz) - gz0;
    D.Nx = M.Nx; D.Ny = M.Ny; D.Nz = M.Nz;
    D.dt = dt;
    D.e0 = e0;
    const size_t esz = S.esz;
    const size_t vol = (size_t)D.ex1 * D.ey1 * D.ez1 * sizeof(double);
    const size_t parts[9] = {vol, vol, M.Nx * sizeof(double), M.Ny * sizeof(double), M.Nz * sizeof(double),
                             M.Nx * esz, M.Ny * esz, M.Nz * esz, sizeof(M.raw)};
    const void* src[9] = {eps_now, sigma, dx, dy, dz, dx1, dy1, dz1, M.raw};
    size_t total = 0;
    for (size_t b : parts) total += align256(b);
    HIP_OK(hipMalloc(&H.buf, total));
    char* d[9];
    char* q = (char*)H.buf;
    for (int k = 0; k < 9; ++k) {
        d[k] = q;
        if (hipMemcpy(q, src[k], parts[k], hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(H.buf);
            RF_THROW(RFDTD_EDEVICE, "design box upload failed");
        }
        q += align256(parts[k]);
    }
    H.eps = (double*)d[0];
    D.eps = H.eps;
    D.sig = (const double*)d[1];
    D.dx = (const double*)d[2];
    D.dy = (const double*)d[3];
    D.dz = (const double*)d[4];
    D.dx1 = d[5];
    D.dy1 = d[6];
    D.dz1 = d[7];
    H.raw = (void* const*)d[8];
    // the tuples of the current permittivity must be what the session holds
    const int blocks = (int)((n + 255) / 256);
    int* d_bad = nullptr;
    HIP_OK(hipMalloc(&d_bad, (size_t)blocks * sizeof(int)));
    if (S.f64)
        k_design_coefficients<double><<<blocks, 256>>>(D, M.coef_kind, (uint32_t*)M.eidx, (double*)M.etab,
                                                       (double* const*)H.raw, 0, 1, d_bad);
    else
        k_design_coefficients<float><<<blocks, 256>>>(D, M.coef_kind, (uint32_t*)M.eidx, (float*)M.etab,
                                                      (float* const*)H.raw, 0, 1, d_bad);
    std::vector<int> bad(blocks);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipMemcpy(bad.data(), d_bad, blocks * sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d_bad);
    bool any = false;
    for (int b : bad) any |= b != 0;
    if (e1 != hipSuccess || e2 != hipSuccess || any) {
        (void)hipFree(H.buf);
        if (e1 != hipSuccess || e2 != hipSuccess) RF_THROW(RFDTD_EDEVICE, "design box check failed to run");
        RF_THROW(RFDTD_EINVAL, "the design box touches edges that are not plain dielectric (PML, conductor, "
                               "lumped element or edge correction), or the current permittivity given does "
                               "not match the session's coefficients");
    }
    if (M.coef_kind == 1) {
        const bool same = M.design_base >= 0 && std::equal(gbox, gbox + 6, M.design_box);
        if (same) {
            H.base = M.design_base;  // the box's cells already own entries [base, base + n)
        } else {
            const int64_t used = M.nE + M.design_used;
            if (M.etab_cap < M.nE) M.etab_cap = M.nE;
            if (used + n > M.etab_cap) {
                const int64_t cap = used + n;
                void* t = nullptr;
                if (hipMalloc(&t, (size_t)cap * 12 * esz) != hipSuccess ||
                    hipMemcpy(t, M.etab, (size_t)used * 12 * esz, hipMemcpyDeviceToDevice) != hipSuccess) {
                    if (t) (void)hipFree(t);
                    (void)hipFree(H.buf);
                    RF_THROW(RFDTD_ENOMEM, "design box entries do not fit in device memory");
                }
                if (M.d_etab_design) (void)hipFree(M.d_etab_design);
                M.d_etab_design = t;
                M.etab = t;
                M.etab_cap = cap;
                S.dirty = true;
                S.graph_dirty = true;
            }
            H.base = used;
            M.design_used += n;
            M.design_base = H.base;
            std::copy(gbox, gbox + 6, M.design_box);
        }
    }
    return H;
}

// The box's 12 E coefficients rewritten from its held permittivity, on stream st.
void design_write(Session& S, Member& M, const DesignHeld& H, hipStream_t st) {
    const DesignDev& D = H.D;
    const int64_t n = (int64_t)D.gnx * D.gny * D.gnz;
    const int blocks = (int)((n + 255) / 256);
    if (S.f64)
        k_design_coefficients<double><<<blocks, 256, 0, st>>>(D, M.coef_kind, (uint32_t*)M.eidx, (double*)M.etab,
                                                              (double* const*)H.raw, H.base, 0, nullptr);
    else
        k_design_coefficients<float><<<blocks, 256, 0, st>>>(D, M.coef_kind, (uint32_t*)M.eidx, (float*)M.etab,
                                                             (float* const*)H.raw, H.base, 0, nullptr);
    HIP_OK(hipGetLastError());
}

}  // namespace

extern "C" {

const char* rfdtd_last_error(void) { return g_last_error.c_str(); }

int32_t rfdtd_version(void) { return 3; }

int rfdtd_create(rfdtd_session** out, int32_t Nx, int32_t Ny, int32_t Nz, const void* const coefficients[30],
                 uint32_t flags) {
    return guarded([&] {
        if (!out) RF_THROW(RFDTD_EINVAL, "null session pointer");
        *out = nullptr;
        *out = (rfdtd_session*)make_session({MemberSpec{Nx, Ny, Nz, 1, coefficients}}, flags);
    });
}

int rfdtd_create_assembled(rfdtd_session** out, const rfdtd_assembly* a, int32_t excitations, uint32_t flags) {
    return guarded([&] {
        if (!out || !a) RF_THROW(RFDTD_EINVAL, "null argument");
        *out = nullptr;
        *out = (rfdtd_session*)make_session({MemberSpec{a->Nx, a->Ny, a->Nz, excitations, nullptr, a}}, flags);
    });
}

int rfdtd_create_batch_assembled(rfdtd_session** out, int32_t n_grids, const rfdtd_assembly* a,
                                 const int32_t* excitations, uint32_t flags) {
    return guarded([&] {
        if (!out || !a || !excitations || n_grids < 1) RF_THROW(RFDTD_EINVAL, "invalid batch arguments");
        *out = nullptr;
        std::vector<MemberSpec> specs;
        for (int g = 0; g < n_grids; ++g)
            specs.push_back(MemberSpec{a[g].Nx, a[g].Ny, a[g].Nz, excitations[g], nullptr, &a[g]});
        *out = (rfdtd_session*)make_session(specs, flags);
    });
}

// The member's packed coefficients over box b (x0, x1, y0, y1, z0, z1) into out[k] (box-sized, C
// order; null skips one), with the x = Nx-1 plane rule applied as the kernels read the arrays.
void read_coefficient_box(Session& s, int member, const int32_t b[6], void* const out[30]) {
    if (member < 0 || member >= (int)s.members.size()) RF_THROW(RFDTD_EINVAL, "member index out of range");
    const Member& M = s.members[member];
    if (b[0] < 0 || b[0] >= b[1] || b[1] > M.Nx || b[2] < 0 || b[2] >= b[3] || b[3] > M.Ny || b[4] < 0 ||
        b[4] >= b[5] || b[5] > M.Nz)
        RF_THROW(RFDTD_EINVAL, "box outside the grid");
    HIP_OK(hipDeviceSynchronize());
    const size_t esz = s.esz;
    const int64_t cells = M.cells;
    const int bx = b[1] - b[0], by = b[3] - b[2], bz = b[5] - b[4];
    auto cell_of = [&](int i, int j, int k) { return ((int64_t)(b[0] + i) * M.Ny + (b[2] + j)) * M.Nz + (b[4] + k); };
    auto put = [&](int k, int64_t q, const char* src) {
        char* v = (char*)out[k] + (size_t)q * esz;
        const int x = (int)(q / ((int64_t)by * bz)) + b[0];
        if (zero_last_plane(k) && x == M.Nx - 1) std::memset(v, 0, esz);
        else std::memcpy(v, src, esz);
    };
    if (M.coef_kind == 0) {
        std::vector<char> row((size_t)bz * esz);
        for (int k = 0; k < 30; ++k) {
            if (!out[k]) continue;
            for (int i = 0; i < bx; ++i)
                for (int j = 0; j < by; ++j) {
                    HIP_OK(hipMemcpy(row.data(), (const char*)M.raw[k] + (size_t)cell_of(i, j, 0) * esz, row.size(),
                                     hipMemcpyDeviceToHost));
                    for (int z = 0; z < bz; ++z) put(k, ((int64_t)i * by + j) * bz + z, &row[(size_t)z * esz]);
                }
        }
        return;
    }
    std::vector<uint32_t> ei(cells), hi(cells);
    HIP_OK(hipMemcpy(ei.data(), M.eidx, (size_t)cells * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(hi.data(), M.hidx, (size_t)cells * 4, hipMemcpyDeviceToHost));
    uint32_t ne = 0, nh = 0;
    for (int i = 0; i < bx; ++i)
        for (int j = 0; j < by; ++j)
            for (int z = 0; z < bz; ++z) {
                const int64_t c = cell_of(i, j, z);
                ne = std::max(ne, ei[c] + 1);
                nh = std::max(nh, hi[c] + 1);
            }
    std::vector<char> et((size_t)ne * 12 * esz), ht((size_t)nh * 20 * esz);
    HIP_OK(hipMemcpy(et.data(), M.etab, et.size(), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ht.data(), M.htab, ht.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < bx; ++i)
        for (int j = 0; j < by; ++j)
            for (int z = 0; z < bz; ++z) {
                const int64_t c = cell_of(i, j, z), q = ((int64_t)i * by + j) * bz + z;
                for (int k = 0; k < 30; ++k) {
                    if (!out[k]) continue;
                    put(k, q, k < 12 ? &et[((size_t)ei[c] * 12 + k) * esz] : &ht[((size_t)hi[c] * 20 + (k - 12)) * esz]);
                }
            }
}

int rfdtd_read_coefficients(rfdtd_session* sp, int32_t member, void* const out[30]) {
    return guarded([&] {
        if (!sp || !out) RF_THROW(RFDTD_EINVAL, "null argument");
        Session& s = *(Session*)sp;
        if (member < 0 || member >= (int)s.members.size()) RF_THROW(RFDTD_EINVAL, "member index out of range");
        const Member& M = s.members[member];
        const int32_t b[6] = {0, M.Nx, 0, M.Ny, 0, M.Nz};
        if (M.coef_kind == 0) {  // whole arrays: one copy each
            HIP_OK(hipDeviceSynchronize());
            const size_t esz = s.esz;
            const int64_t cells = M.cells, plane = (int64_t)M.Ny * M.Nz;
            for (int k = 0; k < 30; ++k) {
                if (!out[k]) continue;
                HIP_OK(hipMemcpy(out[k], M.raw[k], (size_t)cells * esz, hipMemcpyDeviceToHost));
                if (zero_last_plane(k)) std::memset((char*)out[k] + (size_t)(cells - plane) * esz, 0, (size_t)plane * esz);
            }
            return;
        }
        read_coefficient_box(s, member, b, out);
    });
}

int rfdtd_read_coefficient_box(rfdtd_session* sp, int32_t member, const int32_t box[6], void* const out[30]) {
    return guarded([&] {
        if (!sp || !out || !box) RF_THROW(RFDTD_EINVAL, "null argument");
        read_coefficient_box(*(Session*)sp, member, box, out);
    });
}

int rfdtd_create_batch(rfdtd_session** out, int32_t n_grids, const int32_t* dims, const int32_t* excitations,
                       const void* const* coefficients, uint32_t flags) {
    return guarded([&] {
        if (!out || !dims || !excitations || !coefficients || n_grids < 1)
            RF_THROW(RFDTD_EINVAL, "invalid batch arguments");
        *out = nullptr;
        std::vector<MemberSpec> specs;
        for (int g = 0; g < n_grids; ++g)
            specs.push_back(MemberSpec{dims[3 * g], dims[3 * g + 1], dims[3 * g + 2], excitations[g],
                                       coefficients + (size_t)30 * g});
        *out = (rfdtd_session*)make_session(specs, flags);
    });
}

int32_t rfdtd_instance_count(const rfdtd_session* s) { return s ? (int32_t)((const Session*)s)->insts.size() : 0; }

int rfdtd_set_probes(rfdtd_session* s, const rfdtd_probe* probes, int32_t n_probes, int32_t Nt) {
    return guarded([&] { set_probes(S_of(s), 0, probes, n_probes, Nt); });
}

int rfdtd_set_monitors(rfdtd_session* s, const rfdtd_monitor* monitors, int32_t n_monitors, int32_t Nt) {
    return guarded([&] { set_monitors(S_of(s), 0, monitors, n_monitors, Nt); });
}

int rfdtd_set_instance_sources(rfdtd_session* s, int32_t instance, const rfdtd_probe* probes, int32_t n_probes,
                               const rfdtd_monitor* monitors, int32_t n_monitors, int32_t Nt) {
    return guarded([&] {
        set_probes(S_of(s), instance, probes, n_probes, Nt);
        set_monitors(S_of(s), instance, monitors, n_monitors, Nt);
    });
}

int rfdtd_run(rfdtd_session* s, int32_t n_steps, void* stream) {
    return guarded([&] {
        Session& S = S_of(s);
        hipStream_t st = stream ? (hipStream_t)stream : S.own_stream;
        if (S.f64) run_steps<double>(S, n_steps, st); else run_steps<float>(S, n_steps, st);
    });
}

int rfdtd_synchronize(rfdtd_session* s) {
    return guarded([&] {
        Session& S = S_of(s);
        HIP_OK(hipDeviceSynchronize());
        float ms = 0;
        // before any run the events hold no record: the query fails and leaves HIP's last error set,
        // which a later hipGetLastError would report as the failure of an unrelated launch
        if (hipEventElapsedTime(&ms, S.ev0, S.ev1) == hipSuccess) S.last_total_ms = ms;
        else (void)hipGetLastError();
        double se = 0, sh = 0;
        for (int k = 0; k < S.n_timed; ++k) {
            float a = 0, b = 0;
            HIP_OK(hipEventElapsedTime(&a, S.kev[4 * k], S.kev[4 * k + 1]));
            HIP_OK(hipEventElapsedTime(&b, S.kev[4 * k + 2], S.kev[4 * k + 3]));
            se += a;
            sh += b;
        }
        S.last_e_ms = S.n_timed ? se / S.n_timed : 0;
        S.last_h_ms = S.n_timed ? sh / S.n_timed : 0;
    });
}

int rfdtd_read_probes(rfdtd_session* s) {
    return guarded([&] { read_probes(S_of(s)); });
}

int rfdtd_read_monitors(rfdtd_session* s) {
    return guarded([&] { read_monitors(S_of(s)); });
}

int rfdtd_read_fields(rfdtd_session* s, void* const out[18]) {
    return guarded([&] {
        if (!out) RF_THROW(RFDTD_EINVAL, "null argument");
        read_fields(S_of(s), 0, out);
    });
}

int rfdtd_read_instance_fields(rfdtd_session* s, int32_t instance, void* const out[18]) {
    return guarded([&] {
        if (!out) RF_THROW(RFDTD_EINVAL, "null argument");
        read_fields(S_of(s), instance, out);
    });
}

int rfdtd_reset(rfdtd_session* s) {
    return guarded([&] {
        Session& S = S_of(s);
        HIP_OK(hipDeviceSynchronize());
        for (Instance& I : S.insts) {
            HIP_OK(hipMemset(I.d_fields, 0, I.fields_bytes));
            if (I.d_fields2) HIP_OK(hipMemset(I.d_fields2, 0, I.fields_bytes));
            for (auto& h : I.mons) HIP_OK(hipMemset(h.d_values, 0, monitor_bytes(S, h)));
        }
        if (S.copy_stream) HIP_OK(hipStreamSynchronize(S.copy_stream));
        S.cpy_pending[0] = S.cpy_pending[1] = false;
        S.seg_count = 0;
        S.step = 0;
        S.parity = 0;
        S.totals_stale = false;
    });
}

int64_t rfdtd_step(const rfdtd_session* s) { return s ? ((const Session*)s)->step : -1; }

int rfdtd_last_timing(const rfdtd_session* s, double* total_ms, double* e_kernel_ms, double* h_kernel_ms) {
    return guarded([&] {
        if (!s) RF_THROW(RFDTD_EINVAL, "null session");
        const Session& S = *(const Session*)s;
        if (total_ms) *total_ms = S.last_total_ms;
        if (e_kernel_ms) *e_kernel_ms = S.last_e_ms;
        if (h_kernel_ms) *h_kernel_ms = S.last_h_ms;
    });
}

int rfdtd_memory_info(const rfdtd_session* s, uint64_t* device_bytes, int32_t* coef_kind, int64_t* n_e_entries,
                      int64_t* n_h_entries) {
    return guarded([&] {
        if (!s) RF_THROW(RFDTD_EINVAL, "null session");
        const Session& S = *(const Session*)s;
        uint64_t b = 0;
        int64_t ne = 0, nh = 0;
        for (const auto& M : S.members) {
            b += M.coef_bytes;
            ne += M.nE;
            nh += M.nH;
        }
        for (const auto& I : S.insts) b += I.fields_bytes * (I.d_fields2 ? 2 : 1);
        if (device_bytes) *device_bytes = b;
        if (coef_kind) *coef_kind = S.members[0].coef_kind;
        if (n_e_entries) *n_e_entries = ne;
        if (n_h_entries) *n_h_entries = nh;
    });
}

int rfdtd_set_launch(rfdtd_session* s, int32_t block_threads, int32_t x_planes, int32_t graph_steps) {
    return guarded([&] {
        Session& S = S_of(s);
        if (block_threads) {
            if (block_threads < 64 || block_threads > 1024 || block_threads % 64)
                RF_THROW(RFDTD_EINVAL, "block_threads must be a multiple of 64 in [64, 1024]");
            S.block = block_threads;
        }
        if (x_planes) {
            if (x_planes < -1) RF_THROW(RFDTD_EINVAL, "x_planes must be positive, or -1 for automatic");
            for (auto& M : S.members) M.xplanes_req = x_planes < 0 ? 0 : x_planes;
        }
        if (graph_steps) {
            if (graph_steps < 0) RF_THROW(RFDTD_EINVAL, "graph_steps must be positive");
            S.graph_steps = graph_steps;
        }
        S.dirty = true;
        S.graph_dirty = true;
    });
}

int rfdtd_set_variant(rfdtd_session* s, int32_t variant) {
    return guarded([&] {
        Session& S = S_of(s);
        if (variant < 0 || variant > 15) RF_THROW(RFDTD_EINVAL, "variant must be in 0..15");
        if (!vec_fits(S, variant_vec(variant)))
            RF_THROW(RFDTD_EINVAL, "variant " + std::to_string(variant) + " needs Nz a multiple of " +
                                       std::to_string(variant_vec(variant)) + " in every grid");
        apply_variant(S, variant);
        if (!S.fused) release_second_buffers(S);
    });
}

int rfdtd_set_column_order(rfdtd_session* s, int32_t order) {
    return guarded([&] {
        Session& S = S_of(s);
        if (order != 0 && order != 1) RF_THROW(RFDTD_EINVAL, "column order must be 0 (XCD ranges) or 1 (launch order)");
        S.sep_order = order;  // the checkpointed gradient's recomputation too
        if (S.col_order != order) {
            S.col_order = order;
            S.dirty = true;
            S.graph_dirty = true;
        }
    });
}

int rfdtd_get_column_order(rfdtd_session* s, int32_t* order) {
    return guarded([&] {
        if (!order) RF_THROW(RFDTD_EINVAL, "null argument");
        *order = S_of(s).col_order;
    });
}

int rfdtd_set_design_permittivity(rfdtd_session* s, int32_t member, const int32_t gbox[6], const double* eps_now,
                                  const double* eps_new, const double* sigma, const double* dx, const double* dy,
                                  const double* dz, const void* dx1, const void* dy1, const void* dz1, double dt,
                                  double e0) {
    return guarded([&] {
        Session& S = S_of(s);
        if (!eps_new) RF_THROW(RFDTD_EINVAL, "design permittivity arrays missing");
        Member& M = design_member(S, member);
        HIP_OK(hipDeviceSynchronize());
        DesignHeld H = design_hold(S, M, gbox, eps_now, sigma, dx, dy, dz, dx1, dy1, dz1, dt, e0);
        // the new permittivity replaces the held one; the box's coefficients are rewritten from it
        const DesignDev& D = H.D;
        HIP_OK(hipMemcpy(H.eps, eps_new, (size_t)D.ex1 * D.ey1 * D.ez1 * sizeof(double), hipMemcpyHostToDevice));
        design_write(S, M, H, nullptr);
        HIP_OK(hipDeviceSynchronize());
        // a box prepared for device updates now holds other values: it must be prepared again
        if (M.held.buf) (void)hipFree(M.held.buf);
        M.held = Member::DesignHeld{};
        (void)hipFree(H.buf);
    });
}

int rfdtd_design_prepare(rfdtd_session* s, int32_t member, const int32_t gbox[6], const double* eps_now,
                         const double* sigma, const double* dx, const double* dy, const double* dz, const void* dx1,
                         const void* dy1, const void* dz1, double dt, double e0) {
    return guarded([&] {
        Session& S = S_of(s);
        Member& M = design_member(S, member);
        HIP_OK(hipDeviceSynchronize());
        DesignHeld H = design_hold(S, M, gbox, eps_now, sigma, dx, dy, dz, dx1, dy1, dz1, dt, e0);
        if (M.held.buf) (void)hipFree(M.held.buf);
        M.held = H;
        M.held.ready = true;
    });
}

int rfdtd_set_design_permittivity_device(rfdtd_session* s, int32_t member, const int32_t cbox[6], const void* eps_r,
                                         int32_t f64, void* stream) {
    return guarded([&] {
        Session& S = S_of(s);
        Member& M = design_member(S, member);
        if (!M.held.ready) RF_THROW(RFDTD_EINVAL, "no design box prepared on this member (rfdtd_design_prepare)");
        if (!cbox || !eps_r) RF_THROW(RFDTD_EINVAL, "design box or permittivity missing");
        const DesignDev& D = M.held.D;
        // the cells must lie in the held volume: the packed-cell box grown by one cell on each upper side
        const int bx = cbox[0] - D.gx0, by = cbox[2] - D.gy0, bz = cbox[4] - D.gz0;
        const int nx = cbox[1] - cbox[0], ny = cbox[3] - cbox[2], nz = cbox[5] - cbox[4];
        if (nx < 1 || ny < 1 || nz < 1 || bx < 0 || by < 0 || bz < 0 || bx + nx > D.ex1 || by + ny > D.ey1 ||
            bz + nz > D.ez1)
            RF_THROW(RFDTD_EINVAL, "cells outside the prepared design box");
        hipStream_t st = stream ? (hipStream_t)stream : S.own_stream;
        auto order = [](hipStream_t before, hipStream_t after) {  // `after` waits for `before`'s work so far
            hipEvent_t ev;
            HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            HIP_OK(hipEventRecord(ev, before));
            HIP_OK(hipStreamWaitEvent(after, ev, 0));
            (void)hipEventDestroy(ev);
        };
        // the session's earlier steps read the old values: those on its own stream, and the last run,
        // whatever stream it was given (ev1 marks its end; an event never recorded is no wait)
        if (st != S.own_stream) order(S.own_stream, st);
        HIP_OK(hipStreamWaitEvent(st, S.ev1, 0));
        const int64_t n = (int64_t)nx * ny * nz;
        k_design_put_eps<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(M.held.eps, D.ey1, D.ez1, bx, by, bz, nx, ny, nz,
                                                                     eps_r, f64 ? 1 : 0, D.e0);
        HIP_OK(hipGetLastError());
        design_write(S, M, M.held, st);
        // and the session's later steps the new ones: its own stream's work, and the next run on any stream
        if (st != S.own_stream) order(st, S.own_stream);
        if (!S.design_ev) HIP_OK(hipEventCreateWithFlags(&S.design_ev, hipEventDisableTiming));
        HIP_OK(hipEventRecord(S.design_ev, st));
        S.design_pending = true;
    });
}

int rfdtd_host_alloc(uint64_t bytes, void** ptr) {
    return guarded([&] {
        if (!ptr) RF_THROW(RFDTD_EINVAL, "null argument");
        *ptr = nullptr;
        if (bytes == 0) return;
        if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess)
            RF_THROW(RFDTD_ENOMEM, "page-locked host allocation of " + std::to_string(bytes) + " bytes failed");
    });
}

int rfdtd_host_free(void* ptr) {
    return guarded([&] {
        if (ptr) HIP_OK(hipHostFree(ptr));
    });
}

int rfdtd_set_monitor_streaming(rfdtd_session* s, int32_t ring_samples) {
    return guarded([&] {
        Session& S = S_of(s);
        if (ring_samples != 0 && ring_samples < 2) RF_THROW(RFDTD_EINVAL, "ring_samples must be 0 or at least 2");
        S.stream_ring = ring_samples;
    });
}

int rfdtd_set_adjoint_launch(rfdtd_session* s, int32_t variant, int32_t block_threads, int32_t x_chunks) {
    return guarded([&] {
        Session& S = S_of(s);
        if (variant < 0 || variant > 3) RF_THROW(RFDTD_EINVAL, "adjoint variant must be 0, 1, 2 or 3");
        if (block_threads != 0 && (block_threads < 64 || block_threads > 1024 || block_threads % 64))
            RF_THROW(RFDTD_EINVAL, "block_threads must be 0 or a multiple of 64 in [64, 1024]");
        if (x_chunks < 0) RF_THROW(RFDTD_EINVAL, "x_chunks must be >= 0");
        S.adj_pf = variant;
        S.adj_block = block_threads;
        S.adj_chunks = x_chunks;
    });
}

int rfdtd_get_adjoint_launch(rfdtd_session* s, int32_t config_out[3]) {
    return guarded([&] {
        Session& S = S_of(s);
        if (!config_out) RF_THROW(RFDTD_EINVAL, "config_out missing");
        config_out[0] = S.adj_pf;
        config_out[1] = S.adj_block;
        config_out[2] = S.adj_chunks;
    });
}

int rfdtd_autotune(rfdtd_session* s, int32_t trial_steps, int32_t config_out[3], double* ms_per_step) {
    return guarded([&] {
        Session& S = S_of(s);
        if (!config_out) RF_THROW(RFDTD_EINVAL, "config_out missing");
        if (S.f64) autotune<double>(S, trial_steps, config_out, ms_per_step);
        else autotune<float>(S, trial_steps, config_out, ms_per_step);
    });
}

int rfdtd_plan_checkpoints(rfdtd_session* s, int32_t Nt, int32_t n, const int32_t* instances, const int32_t* boxes,
                           uint64_t budget_bytes, int32_t* every_out, uint64_t* bytes_out) {
    return guarded([&] {
        Session& S = S_of(s);
        if (n < 1 || !instances || !boxes || !every_out) RF_THROW(RFDTD_EINVAL, "plan request arrays missing");
        std::vector<GradItem> items(n);
        for (int k = 0; k < n; ++k) {
            const int32_t* b = boxes + 6 * k;
            items[k] = GradItem{instances[k], Box{b[0], b[2], b[4], b[1] - b[0], b[3] - b[2], b[5] - b[4]}, nullptr, 0,
                                nullptr};
        }
        if (S.f64) plan_checkpoints<double>(S, Nt, items, budget_bytes, every_out, bytes_out);
        else plan_checkpoints<float>(S, Nt, items, budget_bytes, every_out, bytes_out);
    });
}

int rfdtd_set_history(rfdtd_session* s, int32_t n, const int32_t* instances, const int32_t* boxes, int32_t n_steps) {
    return guarded([&] {
        Session& S = S_of(s);
        if (n < 0 || (n > 0 && (!instances || !boxes || n_steps < 1)))
            RF_THROW(RFDTD_EINVAL, "history request arrays missing");
        if (n > 0 && S.step != 0) RF_THROW(RFDTD_ESTATE, "record the history from the first step (run after rfdtd_reset)");
        // the same request again (a loop of runs on one session): keep the buffers
        bool same = n == (int)S.hist.size();
        for (int k = 0; same && k < n; ++k)
            same = S.hist[k].inst == instances[k] && S.hist[k].capacity == n_steps &&
                   std::equal(boxes + 6 * k, boxes + 6 * k + 6, S.hist[k].box);
        if (same) return;
        HIP_OK(hipDeviceSynchronize());
        for (auto& h : S.hist)
            if (h.d) (void)hipFree(h.d);
        S.hist.clear();
        for (int k = 0; k < n; ++k) {
            const Instance& I = instance_at(S, instances[k]);
            const Member& M = S.members[I.member];
            const int32_t* b = boxes + 6 * k;
            if (b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] <= b[0] || b[3] <= b[2] || b[5] <= b[4] || b[1] > M.Nx ||
                b[3] > M.Ny || b[5] > M.Nz)
                RF_THROW(RFDTD_EINVAL, "design box outside the grid");
            Session::History h{};
            h.inst = instances[k];
            std::memcpy(h.box, b, sizeof(h.box));
            h.slot_elems = (int64_t)3 * (b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1);
            h.capacity = n_steps;
            const size_t bytes = (size_t)h.slot_elems * n_steps * S.esz;
            if (hipMalloc(&h.d, bytes) != hipSuccess)
                RF_THROW(RFDTD_ENOMEM, "design-box history of " + std::to_string(bytes) + " bytes does not fit");
            S.hist.push_back(h);
        }
        S.dirty = true;
        S.graph_dirty = true;
    });
}

int rfdtd_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
    return guarded([&] {
        size_t f = 0, t = 0;
        HIP_OK(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

int rfdtd_set_checkpointing(rfdtd_session* s, int32_t every) {
    return guarded([&] {
        Session& S = S_of(s);
        if (every < 0) RF_THROW(RFDTD_EINVAL, "checkpoint interval must be >= 0");
        S.ckpt_every = every;
    });
}

int rfdtd_gradient(rfdtd_session* s, int32_t instance, const int32_t box[6], const void* const* adjoint_series,
                   int32_t n_series, double* grad_out) {
    return rfdtd_gradient_batch(s, 1, &instance, box, &adjoint_series, &n_series, &grad_out);
}

int rfdtd_gradient_batch(rfdtd_session* s, int32_t n, const int32_t* instances, const int32_t* boxes,
                         const void* const* const* adjoint_series, const int32_t* n_series, double* const* grad_out) {
    return guarded([&] {
        Session& S = S_of(s);
        if (n < 1 || !instances || !boxes || !adjoint_series || !n_series || !grad_out)
            RF_THROW(RFDTD_EINVAL, "gradient request arrays missing");
        std::vector<GradItem> items(n);
        for (int k = 0; k < n; ++k) {
            const int32_t* b = boxes + 6 * k;
            items[k] = GradItem{instances[k], Box{b[0], b[2], b[4], b[1] - b[0], b[3] - b[2], b[5] - b[4]},
                                adjoint_series[k], n_series[k], grad_out[k]};
        }
        if (S.f64) gradients<double>(S, items);
        else gradients<float>(S, items);
    });
}

void rfdtd_destroy(rfdtd_session* s) {
    if (!s) return;
    delete (Session*)s;
}

int rfdtd_nf2ff(int32_t n_faces, const rfdtd_ff_face* faces, int32_t n_frequencies, const float* beta,
                int32_t n_theta, const float* theta, int32_t n_phi, const float* phi, float eta0, void* out) {
    std::vector<void*> bufs;
    const int rc = guarded([&] {
        if (n_faces < 1 || !faces || n_frequencies < 1 || !beta || n_theta < 1 || !theta || n_phi < 1 || !phi || !out)
            RF_THROW(RFDTD_EINVAL, "nf2ff arguments missing");
        auto up = [&](const void* src, size_t bytes) {
            void* d = nullptr;
            HIP_OK(hipMalloc(&d, std::max<size_t>(bytes, 4)));
            bufs.push_back(d);
            HIP_OK(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
            return d;
        };
        std::vector<FFFace> ff(n_faces);
        int64_t ncells = 0;
        for (int k = 0; k < n_faces; ++k) {
            const rfdtd_ff_face& F = faces[k];
            if (F.axis < 0 || F.axis > 2 || F.n0 < 1 || F.n1 < 1 || !F.r0 || !F.r1 || !F.ds || !F.J || !F.M)
                RF_THROW(RFDTD_EINVAL, "invalid far-field face");
            const size_t plane = (size_t)F.n0 * F.n1;
            ff[k].axis = F.axis;
            ff[k].n0 = F.n0;
            ff[k].n1 = F.n1;
            ff[k].pos = (float)F.position;
            ff[k].cell0 = (int32_t)ncells;
            ff[k].r0 = (const float*)up(F.r0, plane * 4);
            ff[k].r1 = (const float*)up(F.r1, plane * 4);
            ff[k].ds = (const float*)up(F.ds, plane * 4);
            ff[k].J = (const float2*)up(F.J, 3 * (size_t)n_frequencies * plane * 8);
            ff[k].M = (const float2*)up(F.M, 3 * (size_t)n_frequencies * plane * 8);
            ncells += (int64_t)plane;
        }
        if (ncells > INT32_MAX) RF_THROW(RFDTD_EINVAL, "too many far-field cells");
        const FFFace* d_faces = (const FFFace*)up(ff.data(), ff.size() * sizeof(FFFace));
        const float* d_beta = (const float*)up(beta, (size_t)n_frequencies * 4);
        const float* d_theta = (const float*)up(theta, (size_t)n_theta * 4);
        const float* d_phi = (const float*)up(phi, (size_t)n_phi * 4);
        const int64_t ndir = (int64_t)n_theta * n_phi;
        const size_t out_bytes = 2 * (size_t)n_frequencies * ndir * 8;
        float2* d_out = nullptr;
        HIP_OK(hipMalloc(&d_out, out_bytes));
        bufs.push_back(d_out);
        const int64_t blocks = (int64_t)n_frequencies * ((ndir + kFFChunk - 1) / kFFChunk);
        if (blocks > INT32_MAX) RF_THROW(RFDTD_EINVAL, "too many far-field directions");
        k_nf2ff<<<(int)blocks, kFFChunk>>>(d_faces, n_faces, (int)ncells, n_frequencies, d_beta, d_theta, n_theta, d_phi,
                                          n_phi, eta0, d_out);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(out, d_out, out_bytes, hipMemcpyDeviceToHost));
    });
    for (void* p : bufs) (void)hipFree(p);
    return rc;
}

// rfdtd_nf2ff with its inputs staged in, and its result taken from, the static workspace instead of
// device allocations: the same kernel on the same values, so the results are identical.  Requests
// whose staged bytes (each piece rounded up to 256 B) exceed the workspace are refused.
int rfdtd_nf2ff_static(int32_t n_faces, const rfdtd_ff_face* faces, int32_t n_frequencies, const float* beta,
                       int32_t n_theta, const float* theta, int32_t n_phi, const float* phi, float eta0, void* out) {
    return guarded([&] {
        if (n_faces < 1 || !faces || n_frequencies < 1 || !beta || n_theta < 1 || !theta || n_phi < 1 || !phi || !out)
            RF_THROW(RFDTD_EINVAL, "nf2ff arguments missing");
        if (n_faces > kFFMaxFaces) RF_THROW(RFDTD_EINVAL, "nf2ff_static takes at most 64 faces");
        // plan the workspace first: nothing is copied unless everything fits
        double need = 0;  // in double: no overflow however large the counts
        auto round256 = [](double b) { return std::ceil(std::max(b, 4.0) / 256.0) * 256.0; };
        need += round256((double)sizeof(FFFace) * n_faces);
        double ncells = 0;
        for (int k = 0; k < n_faces; ++k) {
            const rfdtd_ff_face& F = faces[k];
            if (F.axis < 0 || F.axis > 2 || F.n0 < 1 || F.n1 < 1 || !F.r0 || !F.r1 || !F.ds || !F.J || !F.M)
                RF_THROW(RFDTD_EINVAL, "invalid far-field face");
            const double plane = (double)F.n0 * (double)F.n1;
            need += 3 * round256(plane * 4) + 2 * round256(3.0 * n_frequencies * plane * 8);
            ncells += plane;
        }
        const double ndir_d = (double)n_theta * (double)n_phi;
        need += round256(4.0 * n_frequencies) + round256(4.0 * n_theta) + round256(4.0 * n_phi);
        need += round256(2.0 * n_frequencies * ndir_d * 8);
        if (need > (double)kFFScratchBytes)
            RF_THROW(RFDTD_EINVAL, "far-field request needs " + std::to_string((long long)need) +
                                       " B of staging, beyond the 64 MiB static workspace");
        if (ncells > INT32_MAX) RF_THROW(RFDTD_EINVAL, "too many far-field cells");
        const int64_t ndir = (int64_t)n_theta * n_phi;
        const int64_t blocks = (int64_t)n_frequencies * ((ndir + kFFChunk - 1) / kFFChunk);
        if (blocks > INT32_MAX) RF_THROW(RFDTD_EINVAL, "too many far-field directions");

        std::lock_guard<std::mutex> lock(g_scratch_mu);
        char* const base = scratch_address();
        size_t used = 0;
        auto take = [&](size_t bytes) {  // next 256-B aligned region of the workspace
            char* p = base + used;
            used += (std::max<size_t>(bytes, 4) + 255) & ~(size_t)255;
            return p;
        };
        auto up = [&](const void* src, size_t bytes) {
            char* d = take(bytes);
            HIP_OK(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
            return (const void*)d;
        };
        char* const d_faces = take(sizeof(FFFace) * n_faces);
        FFFace ff[kFFMaxFaces];
        int64_t cell0 = 0;
        for (int k = 0; k < n_faces; ++k) {
            const rfdtd_ff_face& F = faces[k];
            const size_t plane = (size_t)F.n0 * F.n1;
            ff[k] = FFFace{};
            ff[k].axis = F.axis;
            ff[k].n0 = F.n0;
            ff[k].n1 = F.n1;
            ff[k].pos = (float)F.position;
            ff[k].cell0 = (int32_t)cell0;
            ff[k].r0 = (const float*)up(F.r0, plane * 4);
            ff[k].r1 = (const float*)up(F.r1, plane * 4);
            ff[k].ds = (const float*)up(F.ds, plane * 4);
            ff[k].J = (const float2*)up(F.J, 3 * (size_t)n_frequencies * plane * 8);
            ff[k].M = (const float2*)up(F.M, 3 * (size_t)n_frequencies * plane * 8);
            cell0 += (int64_t)plane;
        }
        HIP_OK(hipMemcpy(d_faces, ff, sizeof(FFFace) * n_faces, hipMemcpyHostToDevice));
        const float* d_beta = (const float*)up(beta, (size_t)n_frequencies * 4);
        const float* d_theta = (const float*)up(theta, (size_t)n_theta * 4);
        const float* d_phi = (const float*)up(phi, (size_t)n_phi * 4);
        const size_t out_bytes = 2 * (size_t)n_frequencies * ndir * 8;
        float2* d_out = (float2*)take(out_bytes);  // every element is written by the kernel
        if (used > kFFScratchBytes) RF_THROW(RFDTD_EINVAL, "far-field workspace plan exceeded");  // not reached
        k_nf2ff<<<(int)blocks, kFFChunk>>>((const FFFace*)d_faces, n_faces, (int)cell0, n_frequencies, d_beta, d_theta,
                                          n_theta, d_phi, n_phi, eta0, d_out);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(out, d_out, out_bytes, hipMemcpyDeviceToHost));
    });
}

int rfdtd_copy_bandwidth(uint64_t bytes, int32_t iters, double* gbytes_per_s, double* ms_per_copy) {
    void* a = nullptr;
    void* b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int rc = guarded([&] {
        if (bytes < 4096 || iters < 1) RF_THROW(RFDTD_EINVAL, "bytes >= 4096 and iters >= 1 required");
        bytes &= ~(uint64_t)15;
        HIP_OK(hipMalloc(&a, bytes));
        HIP_OK(hipMalloc(&b, bytes));
        HIP_OK(hipMemset(a, 0, bytes));
        HIP_OK(hipMemset(b, 0, bytes));
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        const int64_t n = (int64_t)(bytes / 16);
        const int grid = 256 * 32;
        k_stream_copy<<<grid, 256>>>((float4*)b, (const float4*)a, n);  // warm-up
        HIP_OK(hipEventRecord(e0, nullptr));
        for (int k = 0; k < iters; ++k) k_stream_copy<<<grid, 256>>>((float4*)b, (const float4*)a, n);
        HIP_OK(hipEventRecord(e1, nullptr));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        const double per = (double)ms / iters;
        if (ms_per_copy) *ms_per_copy = per;
        if (gbytes_per_s) *gbytes_per_s = 2.0 * (double)bytes / (per * 1e6);
    });
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int rfdtd_grid_barrier_cost(int32_t workgroups, int32_t fences, int32_t iters, double* us_per_barrier,
                            double* us_per_launch) {
    uint32_t* words = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t cs = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    const int rc = guarded([&] {
        if (workgroups < 1 || fences < 0 || fences > 2 || iters < 1 || iters > 100000)
            RF_THROW(RFDTD_EINVAL, "workgroups >= 1, fences 0..2 and iters 1..100000 required");
        int dev = 0, ncu = 0, per_cu = 0, khz = 0, coop = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        HIP_OK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
        const void* kern = fences == 0 ? (const void*)k_barrier_probe<0>
                         : fences == 1 ? (const void*)k_barrier_probe<1> : (const void*)k_barrier_probe<2>;
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
        if (!coop) RF_THROW(RFDTD_EDEVICE, "the device does not support cooperative launches");
        if ((int64_t)workgroups > (int64_t)per_cu * ncu)
            RF_THROW(RFDTD_EINVAL, "more workgroups than can be resident at once (" + std::to_string(per_cu * ncu) + ")");
        const size_t nwords = 64 + 32 * (size_t)workgroups;
        HIP_OK(hipMalloc(&words, nwords * 4));
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        long long limit = (long long)khz * 1000LL;  // one second of wall clock per wait
        int it = iters;
        void* args[] = {&words, &it, &limit};
        float best = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            HIP_OK(hipMemset(words, 0, nwords * 4));
            HIP_OK(hipEventRecord(e0, nullptr));
            HIP_OK(hipLaunchCooperativeKernel(kern, dim3(workgroups), dim3(256), args, 0, nullptr));
            HIP_OK(hipEventRecord(e1, nullptr));
            HIP_OK(hipEventSynchronize(e1));
            uint32_t ab = 0;
            HIP_OK(hipMemcpy(&ab, words + 32, 4, hipMemcpyDeviceToHost));
            if (ab) RF_THROW(RFDTD_EDEVICE, "a barrier wait exceeded its time limit");
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            if (rep == 0 || ms < best) best = ms;
        }
        if (us_per_barrier) *us_per_barrier = 1000.0 * best / iters;
        // the same number of dependent, graph-replayed launches of an empty kernel of the same grid
        constexpr int kLaunches = 100;
        HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kLaunches; ++k) k_boundary_probe<<<workgroups, 256, 0, cs>>>();
        HIP_OK(hipStreamEndCapture(cs, &g));
        HIP_OK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        HIP_OK(hipGraphLaunch(ge, nullptr));
        const int reps = std::max(1, iters / kLaunches);
        HIP_OK(hipEventRecord(e0, nullptr));
        for (int r = 0; r < reps; ++r) HIP_OK(hipGraphLaunch(ge, nullptr));
        HIP_OK(hipEventRecord(e1, nullptr));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        if (us_per_launch) *us_per_launch = 1000.0 * ms / (reps * kLaunches);
    });
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (cs) (void)hipStreamDestroy(cs);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (words) (void)hipFree(words);
    return rc;
}

int rfdtd_solver_run(const void* const coefficients[30], const rfdtd_probe* probes, int32_t n_probes,
                     const rfdtd_monitor* monitors, int32_t n_monitors, int32_t Nx, int32_t Ny, int32_t Nz,
                     int32_t Nt, int32_t update_interval, uint32_t flags) {
    Session* s = nullptr;
    const int rc = guarded([&] {
        if (Nt < 1) RF_THROW(RFDTD_EINVAL, "Nt must be positive");
        s = make_session({MemberSpec{Nx, Ny, Nz, 1, coefficients}}, flags);
        set_probes(*s, 0, probes, n_probes, Nt);
        set_monitors(*s, 0, monitors, n_monitors, Nt);
        const bool show = update_interval > 0 && !(flags & RFDTD_QUIET);
        if (show) progress_bar(0, Nt);
        // progress is reported per chunk of update_interval steps, as the reference's controller does
        const int chunk = show ? std::max<int>(update_interval, s->graph_steps) : Nt;
        for (int done = 0; done < Nt;) {
            const int k = std::min(chunk, Nt - done);
            if (s->f64) run_steps<double>(*s, k, s->own_stream); else run_steps<float>(*s, k, s->own_stream);
            done += k;
            if (show) {
                HIP_OK(hipStreamSynchronize(s->own_stream));
                progress_bar(done - 1, Nt);
            }
        }
        HIP_OK(hipStreamSynchronize(s->own_stream));
        if (show) progress_bar(Nt, Nt);
        read_probes(*s);
        read_monitors(*s);
    });
    delete s;
    return rc;
}

}  // extern "C"
