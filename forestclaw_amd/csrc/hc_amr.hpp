// hc_amr.hpp -- device data structures of the AMR driver: ghost-fill
// operations, physical-BC and time-interpolation lists, regrid transfers.
#pragma once

namespace hc {

constexpr int MAXL = 22;  // levels 0..21

// Per-level current solution buffers plus the shared time-interpolated pool.
// A slot index addresses the same tile position in every pool.
template <class R> struct Pools {
    R* q[MAXL];   // current q of level L (one of two ping-pong pools)
    R* qlast[MAXL];
    R* ti;        // time-interpolated data (coarse level of a subcycle)
};

enum GhostKind {
    G_COPY_FACE = 0,     // fclaw2d_clawpatch46_fort_copy_face
    G_COPY_CORNER = 1,   // fclaw2d_clawpatch46_fort_copy_corner
    G_AVG_FACE = 2,      // fclaw2d_clawpatch46_fort_average_face
    G_AVG_CORNER = 3,    // fclaw2d_clawpatch46_fort_average_corner
    G_INTERP_FACE = 4,   // fclaw2d_clawpatch46_fort_interpolate_face
    G_INTERP_CORNER = 5  // fclaw2d_clawpatch46_fort_interpolate_corner
};

// One patch-pair exchange.  For COPY/AVG the destination is the patch whose
// ghost cells are written (this patch / coarse patch) and sub its face or
// corner; for INTERP the destination is the fine patch and sub is the face
// or corner of the coarse (source) patch.  (ox, oy) is the index offset
// between the two tiles: same size, src = dst + (ox, oy); across levels,
// fine = 2*coarse - mbc + (ox, oy) in 0-based tile indices.
struct GhostOp {
    int kind, sub;
    int dst, src;       // slots
    int dlev, slev;     // levels (select pools)
    int ox, oy;
    int flags;          // bit0: dst uses the time-interpolated pool, bit1: src does
};

struct BcOp {           // physical boundary conditions on one patch
    int slot, level, flags;  // flags bit0: time-interpolated pool
    int mthbc[4];
};

struct TiOp {           // time interpolation of one coarse patch
    int slot, level;
};

// Conservative time-sync corrections of one patch (fclaw2d_time_sync.c:118-154
// through cb_face_fill's TIME_SYNC_SAMESIZE / TIME_SYNC_F2C dispatch,
// src/fclaw2d_face_neighbors.c:320-390).
enum TsKind { TS_NONE = 0, TS_SAME = 1, TS_FINE = 2 };
struct TsOp {
    int slot;
    int kind[4];        // per face
    int nbr[4][2];      // same size: [0]; finer: igrid 0, 1
};

// The last phase of a ghost fill (physical BCs) and the time sync of the same
// levels on one patch, run by one workgroup in that order: the BCs read the
// patch's interior cells before the corrections change its edge cells, as
// when the sync follows the fill.  bc.slot < 0: no physical boundary;
// ts.slot < 0: no correction.
struct BcTsOp {
    BcOp bc;
    TsOp ts;
    int level;
};

// The Courant-number reduction that ends a coarse step (the maximum of the n
// slots acc[0..n) into *out, slots cleared), run by one extra workgroup of the
// coarse step's last ghost-fill launch; acc == nullptr: none.
struct CflFin {
    unsigned long long* acc;
    int n;
    double* out;
};

// Several operation lists in one launch: list s holds the operations
// [start[s], start[s+1]) of the launch (one list per level of a ghost-fill
// phase or a time sync, whose levels are independent of each other).
#if defined(__HIPCC__)
#define HC_HD __host__ __device__
#else
#define HC_HD
#endif
constexpr int SEG_MAX = MAXL + 1;
template <class OP, int CAP = SEG_MAX> struct Segs {
    const OP* ops[CAP];
    int start[CAP + 1];
    int n;
    bool ok;            // false once a list did not fit (the launch is refused)
    HC_HD Segs() : n(0), ok(true) { start[0] = 0; }
    // host: append a list (empty lists are skipped)
    void add(const OP* p, int cnt)
    {
        if (cnt <= 0 || !p) return;
        if (n >= CAP) { ok = false; return; }
        ops[n] = p;
        start[n + 1] = start[n] + cnt;
        n++;
    }
    HC_HD int total() const { return start[n]; }
#ifdef __HIPCC__
    // device: list holding operation o of the launch (o < total())
    __device__ int seg(int o) const
    {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (o >= start[mid]) lo = mid; else hi = mid - 1;
        }
        return lo;
    }
    __device__ const OP& at(int o) const
    {
        const int s = seg(o);
        return ops[s][o - start[s]];
    }
#endif
};
// register resets of a time sync: three face lists per level
using ResetSegs = Segs<int, 3 * MAXL>;
// time interpolations run in the average launch of a ghost fill: the weight
// of each list (one list per interpolated level), and per level the weight of
// an interpolation pending in the same launch (bit L of pend set), for the
// copies that read the interpolated pool of that level in that launch
template <class R> struct TiWeights {
    R a[SEG_MAX];
    R al[MAXL];
    unsigned pend;
};

struct XferOp {         // regrid transfer into the new pools
    int kind;           // 0 keep (copy), 1 interpolate2fine, 2 average2coarse
    int dst;            // new slot
    int src[4];         // old slots (keep/refine: [0])
    int srclev;         // level of the old data (selects its current pool)
    int igrid;
};

}  // namespace hc
