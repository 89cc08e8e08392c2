// sph_kernels.hip -- gfx950 kernels of the smooth-particle hot path.
//
// Every pass is a per-particle gather: work-item s owns slot s and sums over
// its own neighbour list, so each pair is evaluated from both ends and no
// float atomics are needed.  Lists are built from a periodic cell list with
// a skin and reused until the particles of a box could have moved more than
// half the skin; slots are sorted by (box, cell) at each build, so the 64
// lanes of a wavefront own a few adjacent cells and their neighbour reads hit
// the same cache lines.
//
// Passes (reference functions they restate in brackets):
//   build  bin + stable radix sort + permute + neighbour lists
//                                      [VerletList.build, neighbour_list.py:160-189]
//   P1     density (h and h_l) + vdW EOS [properties.py:63-120, f_properties.py:100-120]
//   P2     velocity gradient + heat flux [f_properties.py:111,130 (Fortran absent)]
//   P3     pressure/cohesion/viscous/conduction rates
//                                      [forces.py:338-405, c_forces.pyx:96-237,
//                                       spam_complete_force.py:158-165]
//   P4     leapfrog kick/drift, predictor, periodic wrap, skin check
#include <math.h>
#include <type_traits>
#include "pyticles_hip.h"
#include "sph_common.hpp"
#include "sph_kernels.hpp"
#include "sph_sort.hpp"


namespace sph {

static constexpr int NT = 256;  // work-items per block (4 waves)
// The pair passes take their block size at launch (ArenaView::pass_block,
// a schedule the autotuner picks): 128, 256 or 512 work-items.
static constexpr int PASS_NT_MAX = 512;
static constexpr double SPH_PI = 3.14159265358979323846;  // math.pi (spkernel.py:15)
// The pair passes are templates on the number of list entries whose loads
// are in flight per lane (U64 / U32 = 1, 2 or 4); the schedule picks one at
// launch time.  Defaults: 2 for FP64, whose wider registers cut occupancy,
// 4 for FP32.

// gfx950 dispatches block i to XCD (i mod 8).  Remap so that each XCD gets a
// contiguous run of blocks: neighbouring slots (neighbouring cells) then share
// that XCD's L2 instead of being spread over all eight.
__device__ __forceinline__ int xcd_block()
{
    const int G = gridDim.x, pid = blockIdx.x;
    const int x = pid & 7, local = pid >> 3;
    const int per = G >> 3, rem = G & 7;
    return x < rem ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

__device__ __forceinline__ int wrapi(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

// Wave-uniform values: a value the lanes agree on, moved to a scalar
// register (readfirstlane).
__device__ __forceinline__ double wave_uniform(double x)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ float wave_uniform(float x)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Neighbour lists in tiles of 64 slots (one wavefront), entries in groups
// of four.  Short entry k sits at position P = k, long entry k at
// P = lcap - 1 - k (lcap a multiple of 8).  A tile is lcap / 4 rows of 64
// 16-byte cells, cell (h, s % 64) belonging to slot s, in one of two forms:
//  * 32-bit: the neighbour's slot, position P in cell P / 4, word P % 4
//    (word ((s / 64) * (lcap / 4) + P / 4) * 256 + (s % 64) * 4 + P % 4);
//  * 16-bit (the default): the neighbour's slot offset from the slot
//    itself, taken modulo the box's slot count n when the plain difference
//    does not fit (cnt[s].x bit 31 marks a slot holding such an entry),
//    position P in cell P / 8, halfword P % 8, so a slot's entries fill
//    the first lcap / 8 of its cells.
// Either way a slot writes and reads its own cells only, so boxes of both
// forms may share a tile.
// Slots are in cell order, so a neighbour lies within about two rows of
// cells of the slot in slot order, or as far from the box's other end (the
// periodic image): offsets fit 16 bits unless two rows of cells hold more
// than ~32k particles.  A build that meets a neighbour whose offset does not
// fit holds its box for that step (BoxDyn::relist) and the runtime switches
// the box to the 32-bit form, with the same entries in the same order.
// The walks fetch up to four entries per lane with one load and the build
// stores four at a time.
static constexpr int LIST_TILE = 64;
static constexpr int LIST_GROUP = 4;
static constexpr int LIST_ROW = LIST_TILE * LIST_GROUP;  // entries per group row of a tile
static constexpr int CNT_WRAP = (int)0x80000000;       // cnt.x: a 16-bit entry is taken modulo n
static constexpr int CNT_MASK = 0x7fffffff;

__device__ __forceinline__ size_t list_tile_words(const ArenaView &A, int s)
{
    return (size_t)(s / LIST_TILE) * (A.lcap / LIST_GROUP) * LIST_ROW;
}

__device__ __forceinline__ unsigned *list_base(const ArenaView &A, int s)
{
    return A.nbr + list_tile_words(A, s) + (s % LIST_TILE) * LIST_GROUP;
}

__device__ __forceinline__ unsigned short *list_base16(const ArenaView &A, int s)
{
    return (unsigned short *)(A.nbr + list_tile_words(A, s)) + (s % LIST_TILE) * (2 * LIST_GROUP);
}

// Halfword of position P in a slot's 16-bit list (list_base16): cell P / 8
// (cells CELL16 halfwords apart: consecutive rows of the tile), halfword P % 8.
// (Cells on every other row, spread over the tile as the 32-bit form's,
// measured the same.)
static constexpr int CELL16 = 2 * LIST_ROW;
__device__ __forceinline__ size_t pos16(int P) { return (size_t)(P >> 3) * CELL16 + (P & 7); }

// Entries at positions P0 .. P0 + U - 1 of a slot (U = 1, 2 or 4, P0 a
// multiple of U): one load.  32-bit form: one entry per word of w.
template <int U>
__device__ __forceinline__ void load_entries(const unsigned *base, int P0, unsigned (&w)[U])
{
    const unsigned *p = base + (size_t)(P0 / LIST_GROUP) * LIST_ROW + (P0 % LIST_GROUP);
    if constexpr (U == 4) {
        const uint4 g = *(const uint4 *)p;
        w[0] = g.x; w[1] = g.y; w[2] = g.z; w[3] = g.w;
    } else if constexpr (U == 2) {
        const uint2 g = *(const uint2 *)p;
        w[0] = g.x; w[1] = g.y;
    } else {
        w[0] = *p;
    }
}

// 16-bit form: entry i of the U in the low (i even) or high (i odd) half
// of w[i / 2].
template <int U>
__device__ __forceinline__ void load_entries16(const unsigned short *base, int P0, unsigned (&w)[U])
{
    const unsigned short *p = base + pos16(P0);
    if constexpr (U == 4) {
        const uint2 g = *(const uint2 *)p;
        w[0] = g.x; w[1] = g.y;
    } else if constexpr (U == 2) {
        w[0] = *(const unsigned *)p;
    } else {
        w[0] = *p;
    }
}

// Entry form of a wave's lists, one of
//   LF32: 32-bit slots;
//   LF16: 16-bit offsets whose plain sum with the slot is the neighbour;
//   LF16W: 16-bit offsets of which some lane's are taken modulo n (the
//          sum is wrapped into the box's slot range [lo, lo + n)).
// The walks are instantiated per form; list_fmt picks the wave's.
enum : int { LF32 = 0, LF16 = 1, LF16W = 2 };

struct ListFmt {
    int lo, hi, nb;  // the box's slot range [lo, hi), nb = hi - lo
    int form;
};

// The lists of the box bu that the wave's active lanes belong to.
__device__ __forceinline__ ListFmt list_fmt(const ArenaView &A, int bu, int s)
{
    ListFmt f;
    f.lo = wave_uniform(A.boxes[bu].part_off);
    f.nb = wave_uniform(A.boxes[bu].n);
    f.hi = f.lo + f.nb;
    const bool w16 = wave_uniform(A.dyn[bu].list16) != 0;
    f.form = !w16 ? LF32 : (__ballot(A.cnt[s].x < 0) != 0ull ? LF16W : LF16);
    return f;
}

// Calls f with the wave's form as a compile-time constant.
template <class Fn>
__device__ __forceinline__ void with_form(const ListFmt &L, Fn f)
{
    if (L.form == LF16) f(std::integral_constant<int, LF16>());
    else if (L.form == LF16W) f(std::integral_constant<int, LF16W>());
    else f(std::integral_constant<int, LF32>());
}

// A slot's short (LONG = false) or long entries, U per round (k0 a multiple
// of U; long entry k sits at position lcap - 1 - k), in list form LF.
// take(k0, t) gives the slots of round k0.  With AHEAD, next(k0) issues the
// load of the next round's group, so that list fetch is in flight while
// round k0's neighbour records are gathered and summed; at the last round it
// re-reads the current group, so every load stays inside the slot's list
// rows.  Without AHEAD, take() loads its own round's group.  Entries past the
// count read as s (a valid record whose terms the callers mask off), so the
// gathers need no branch.  AHEAD pays on the density walks (~15 rounds per
// slot); on the short walks of the gradient and rate passes (3-4 rounds) its
// extra load and registers cost more than the overlap gains.
template <int U, bool LONG, bool AHEAD, int LF>
struct ListRounds {
    // 16-bit entries four per round: one 16-byte cell (8 entries) serves a
    // pair of rounds, so a cell load every other round (issued, with AHEAD,
    // two rounds before its use); c holds the current cell
    static constexpr bool PAIRS = LF != LF32 && U == 4;
    const unsigned *base;
    const unsigned short *base16;
    int lcap, n, s, lo, hi, nb;
    unsigned w[U];
    unsigned c[PAIRS ? 4 : 1];
    __device__ __forceinline__ ListRounds(const ArenaView &A, int s_, int n_, const ListFmt &L)
        : lcap(A.lcap), n(n_), s(s_), lo(L.lo), hi(L.hi), nb(L.nb)
    {
        if (LF == LF32) base = list_base(A, s_);
        else base16 = list_base16(A, s_);
        if (AHEAD) {
            if constexpr (PAIRS) load_cell(cell(0), w);
            else load(pos(0));
        }
    }
    __device__ __forceinline__ int pos(int k0) const { return LONG ? lcap - k0 - U : k0; }
    // the cell of round k0's entries (lcap a multiple of 8)
    __device__ __forceinline__ int cell(int k0) const { return LONG ? (lcap >> 3) - 1 - (k0 >> 3) : (k0 >> 3); }
    __device__ __forceinline__ void load_cell(int q, unsigned (&d)[4]) const
    {
        const uint4 g = *(const uint4 *)(base16 + (size_t)q * CELL16);
        d[0] = g.x; d[1] = g.y; d[2] = g.z; d[3] = g.w;
    }
    __device__ __forceinline__ void load(int P0)
    {
        if constexpr (LF == LF32) load_entries<U>(base, P0, w);
        else load_entries16<U>(base16, P0, w);
    }
    __device__ __forceinline__ int decode(unsigned x, int half) const
    {
        int e = s + (half ? ((int)x >> 16) : ((int)(x << 16) >> 16));
        if constexpr (LF == LF16W) {
            e = e < lo ? e + nb : e;
            e = e >= hi ? e - nb : e;
        }
        return e;
    }
    __device__ __forceinline__ void take(int k0, int (&t)[U])
    {
        if constexpr (PAIRS) {
            if ((k0 & 4) == 0) {  // first round of a cell
                if constexpr (AHEAD) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) c[i] = w[i];
                } else {
                    load_cell(cell(k0), c);
                }
            }
            // short rounds take offsets 0-3 then 4-7 of the cell, long
            // rounds 7-4 then 3-0 (the half picked by selects: a register
            // array indexed at run time would live in scratch memory)
            const bool hi4 = LONG ? ((k0 & 4) == 0) : ((k0 & 4) != 0);
            const unsigned x0 = hi4 ? c[2] : c[0], x1 = hi4 ? c[3] : c[1];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int o = LONG ? 3 - u : u;  // offset within the half
                t[u] = k0 + u < n ? decode(o < 2 ? x0 : x1, o & 1) : s;
            }
            return;
        } else {
            if (!AHEAD) load(pos(k0));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = LONG ? U - 1 - u : u;
                int e;
                if constexpr (LF == LF32) e = (int)w[i];
                else e = decode(w[i >> 1], i & 1);
                t[u] = k0 + u < n ? e : s;
            }
        }
    }
    // issued after round k0's gathers, so waiting for them does not wait for
    // this load too (loads retire in order); at the last round it re-reads
    // the current group or cell, so every load stays inside the slot's rows
    __device__ __forceinline__ void next(int k0)
    {
        if constexpr (AHEAD) {
            if constexpr (PAIRS) {
                if ((k0 & 4) == 0) load_cell(cell(k0 + 8 < n ? k0 + 8 : k0), w);
            } else {
                load(pos(k0 + U < n ? k0 + U : k0));
            }
        }
    }
};

// Record j of array p through a 32-bit byte offset from the array's base,
// which the walks hold in scalar registers: the gather is then a
// global_load with a scalar base and a 32-bit vector offset, one shift per
// array, where a 64-bit index took a sign extension, a 64-bit shift and a
// 64-bit add.  Slot numbers stay below 2^32 / sizeof(T): the arena holds
// at most 2^27 slots, and the largest record gathered is 32 bytes.
template <class T>
__device__ __forceinline__ const T &at32(const T *p, int j)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + (uint32_t)j * (uint32_t)sizeof(T));
}

// Minimum image exactly as neighbour_list.py:110-123 (all '>' tests, then '<').
template <class Box>
__device__ __forceinline__ void min_image(double &dx, double &dy, const Box &B)
{
    dx = dx > B.hLx ? dx - B.Lx : dx;
    dy = dy > B.hLy ? dy - B.Ly : dy;
    dx = dx < -B.hLx ? dx + B.Lx : dx;
    dy = dy < -B.hLy ? dy + B.Ly : dy;
}

// rsq = dx*dx + dy*dy with no fma, so the strict neighbour test is bit-exact
// against a plain IEEE evaluation of the reference's expression.
__device__ __forceinline__ double rsq_exact(double dx, double dy)
{
#pragma clang fp contract(off)
    const double xx = dx * dx;
    const double yy = dy * dy;
    return xx + yy;
}

__device__ __forceinline__ double wrap_coord(double x, double L)
{
    if (x >= L) x -= L;
    else if (x < 0.0) x += L;
    return x;
}

// Calls f(t0, t1, ox, oy) for the slot range of every stencil cell.
template <class F>
__device__ __forceinline__ void for_each_cell(const BoxDesc &B, const int *__restrict__ cs, int cell, F &&f)
{
    const int cx = cell % B.ncx, cy = cell / B.ncx;
    for (int oy = -B.sy; oy <= B.sy; ++oy) {
        const int yy = wrapi(cy + oy, B.ncy);
        const int *row = cs + B.key_base + yy * B.ncx;
        for (int ox = -B.sx; ox <= B.sx; ++ox) {
            const int xx = wrapi(cx + ox, B.ncx);
            f(row[xx], row[xx + 1], ox, oy);
        }
    }
}

__device__ __forceinline__ bool state_in(int st, int mask) { return ((1 << st) & mask) != 0; }

__device__ __forceinline__ void flag_health(BoxDyn *dyn, int b, int bits) { atomicOr(&dyn[b].health, bits); }

// A box rebuilds its cell order and neighbour list in this launch when it is
// initialising, when its running schedule says so, or when forced (the
// stand-alone passes build every box whose state is in force_mask).
__device__ __forceinline__ bool building(const BoxDyn &D, int force_mask)
{
    if (state_in(D.state, force_mask)) return true;
    return D.state == BOX_INIT || (D.state == BOX_RUN && D.k >= D.next_rebuild);
}

// atomicMax on a non-negative double through its bit pattern (monotone for
// non-negative values); reads first so that most lanes issue no atomic.
__device__ __forceinline__ void max_bits(unsigned long long *addr, double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    if (b > *(volatile unsigned long long *)addr) atomicMax(addr, b);
}

// FP32C neighbour record of a particle: its position as 48-bit box-relative
// fixed point (2^48 units per side), split into the high 32 bits, which go
// with the predicted velocity into one 16-byte record rf = (X_hi, Y_hi, vx,
// vy), and the low 16 bits of each axis, packed into rlo.  The 32-bit
// difference of two X_hi, read as a signed int, is the minimum-image
// separation in units of side / 2^32 (the FP32 long-range walk's
// separation: 1.2e-7 at a side of 512, about one FP32 rounding); with the
// low bits the 48-bit difference is the minimum-image separation of the two
// recorded positions exactly (units of side / 2^48: 1.8e-12 at 512), which
// the short-range terms take in FP64.  Coordinates are in [0, L); one that
// rounds up to L wraps to 0, as the box does.
static constexpr double FIX48 = 281474976710656.0;  // 2^48
static constexpr double FIX32 = 4294967296.0;       // 2^32

__device__ __forceinline__ void fix_rec(double2 r, double2 v, const BoxDesc &B, float4 &hi, unsigned &lo)
{
    const unsigned long long M = (1ull << 48) - 1ull;
    const unsigned long long X = (unsigned long long)fmax(rint(r.x * (FIX48 / B.Lx)), 0.0) & M;
    const unsigned long long Y = (unsigned long long)fmax(rint(r.y * (FIX48 / B.Ly)), 0.0) & M;
    hi = make_float4(__uint_as_float((unsigned)(X >> 16)), __uint_as_float((unsigned)(Y >> 16)), (float)v.x,
                     (float)v.y);
    lo = (unsigned)(X & 0xffffull) | ((unsigned)(Y & 0xffffull) << 16);
}

// Runs f(bu) once for each distinct box bu among the active lanes, on the
// lanes of that box, with bu wave-uniform (a waterfall over the boxes of the
// wave: one round for all waves but those straddling a box boundary).
template <class F>
__device__ __forceinline__ void per_box(int b, F &&f)
{
    bool done = false;
    for (;;) {
        if (!done) {
            const int bu = __builtin_amdgcn_readfirstlane(b);
            if (b == bu) {
                f(bu);
                done = true;
            }
        }
        if (__ballot(!done) == 0ull) break;
    }
}

// ===========================================================================
// build: binning, permutation, cell starts, neighbour lists
// ===========================================================================

// The build stage sorts the building boxes only.  A plan numbers their slots
// consecutively, box after box in slot order (box b's first slot gets the
// compact position box_coff[b]); the radix sort runs over those M compact
// positions, so boxes that keep their lists cost it nothing.  Keys are global
// (box key base + cell, box key ranges increase with the slot offset), so
// after a stable sort the j-th sorted element belongs at the j-th building
// slot.  The other kernels run over slots (a wave = one 64-slot tile) and
// skip slots of boxes that do not build.

static constexpr int PLAN_THREADS = 1024;
static constexpr int PLAN_SEGS = 4096;  // segment tables up to this size are cached in LDS

// plan_n[1] = M (building slots); box_coff[b] for each building box b.
// Building tiles of segment g (64-slot tiles holding its slots, if its box
// builds): count, first tile in t0.  A tile is listed by the segment holding
// its first building slot, so a tile shared with an earlier building
// segment is left to that one.
__device__ __forceinline__ int seg_tiles(const ArenaView &A, int g, int force_mask, int &t0)
{
    const SegDesc sg = A.segs[g];
    if (sg.box < 0 || sg.len <= 0 || !building(A.dyn[sg.box], force_mask)) return 0;
    t0 = sg.off >> 6;
    const int t1 = (sg.off + sg.len - 1) >> 6;
    for (int k = g - 1; k >= 0; --k) {
        const SegDesc pv = A.segs[k];
        if (pv.off + pv.len <= t0 * 64) break;  // ends before tile t0
        if (pv.box >= 0 && pv.len > 0 && building(A.dyn[pv.box], force_mask)) {
            ++t0;
            break;
        }
    }
    return t1 >= t0 ? t1 - t0 + 1 : 0;
}

// Inclusive scan of x over the block into ss[tid]: shuffle scans within
// the waves, then one wave scans the wave totals (three barriers; a
// Hillis-Steele scan in LDS took twenty).
__device__ __forceinline__ void plan_scan(int x, int *ss, int *wsum)
{
    constexpr int NW = PLAN_THREADS / 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        int v = lane < NW ? wsum[lane] : 0;
#pragma unroll
        for (int off = 1; off < NW; off <<= 1) {
            const int y = __shfl_up(v, off);
            if (lane >= off) v += y;
        }
        if (lane < NW) wsum[lane] = v;
    }
    __syncthreads();
    ss[tid] = x + (w > 0 ? wsum[w - 1] : 0);
    __syncthreads();
}

__global__ __launch_bounds__(PLAN_THREADS) void k_build_plan(ArenaView A, int force_mask)
{
    __shared__ int ss[PLAN_THREADS];
    const int tid = threadIdx.x;
    const int per = (A.nseg + PLAN_THREADS - 1) / PLAN_THREADS;
    const int g0 = tid * per, g1 = min(g0 + per, A.nseg);
    int sz = 0;
    for (int g = g0; g < g1; ++g) {
        const SegDesc sg = A.segs[g];
        if (sg.box >= 0 && building(A.dyn[sg.box], force_mask)) sz += sg.len;
    }
    __shared__ int wsum[PLAN_THREADS / 64];
    plan_scan(sz, ss, wsum);
    int pos = ss[tid] - sz;
    for (int g = g0; g < g1; ++g) {
        const SegDesc sg = A.segs[g];
        if (sg.box >= 0 && building(A.dyn[sg.box], force_mask)) {
            A.box_coff[sg.box] = pos;
            pos += sg.len;
            // the shared waves' part of this build's maxima (build_tail)
            // starts from zero, as the whole waves' wbuild entries do
            A.dyn[sg.box].vmax2 = 0ull;
            A.dyn[sg.box].amax2 = 0ull;
        }
    }
    if (tid == PLAN_THREADS - 1) A.plan_n[1] = ss[tid];
    // the building tiles, compacted in slot order
    __shared__ int seg_t0[PLAN_SEGS], seg_end[PLAN_SEGS];  // per segment: first tile, inclusive count
    const bool cached = A.nseg <= PLAN_SEGS;
    __syncthreads();
    int nt = 0;
    for (int g = g0; g < g1; ++g) {
        int t0 = 0;
        const int c = seg_tiles(A, g, force_mask, t0);
        if (cached) {
            seg_t0[g] = t0;
            seg_end[g] = c;  // the range's running count, offset below
        }
        nt += c;
    }
    plan_scan(nt, ss, wsum);
    const int T = ss[PLAN_THREADS - 1];
    if (cached) {
        int run = ss[tid] - nt;
        for (int g = g0; g < g1; ++g) {
            run += seg_end[g];
            seg_end[g] = run;
        }
        __syncthreads();
        // entry i: the segment whose inclusive count first passes i
        for (int i = tid; i < T; i += PLAN_THREADS) {
            int lo = 0, hi = A.nseg - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (seg_end[mid] > i) hi = mid;
                else lo = mid + 1;
            }
            const int before = lo > 0 ? seg_end[lo - 1] : 0;
            A.tile_list[i] = seg_t0[lo] + (i - before);
        }
        if (tid == PLAN_THREADS - 1) A.plan_n[2] = T;
        return;
    }
    // more segments than the cache holds: entry i of the list, written by
    // thread i % PLAN_THREADS (one large box's tiles spread over the block):
    // the thread range holding it by a binary search over the inclusive
    // counts, then a walk over that range's segments
    for (int i = tid; i < T; i += PLAN_THREADS) {
        int lo = 0, hi = PLAN_THREADS - 1;  // first range whose count passes i
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ss[mid] > i) hi = mid;
            else lo = mid + 1;
        }
        int base = lo > 0 ? ss[lo - 1] : 0;
        const int ga = lo * per, gb = min(ga + per, A.nseg);
        for (int g = ga; g < gb; ++g) {
            int t0 = 0;
            const int c = seg_tiles(A, g, force_mask, t0);
            if (i < base + c) {
                A.tile_list[i] = t0 + (i - base);
                break;
            }
            base += c;
        }
    }
    if (tid == PLAN_THREADS - 1) A.plan_n[2] = T;
}

// Box of slot s if it builds in this launch, else -1.
__device__ __forceinline__ int building_box(const ArenaView &A, int s, int force_mask)
{
    const int b = A.slot_box[s];
    return (b >= 0 && building(A.dyn[b], force_mask)) ? b : -1;
}

// Slot owners, written when the layout changes (free slots: -1).
__global__ __launch_bounds__(NT) void k_slot_box(ArenaView A)
{
    const int s = blockIdx.x * NT + threadIdx.x;
    if (s >= A.n) return;
    int lo = 0, hi = A.nseg - 1;
    while (lo < hi) {  // last segment with off <= s
        const int mid = (lo + hi + 1) >> 1;
        if (A.segs[mid].off <= s) lo = mid;
        else hi = mid - 1;
    }
    A.slot_box[s] = A.segs[lo].box;
}

// key = box key base + current cell (row major) at the slot's compact
// position.
// The build kernels run over the building tiles of the plan
// (tile_list[0 .. plan_n[2])), one wave per tile.  Blocks are dealt to the
// XCDs in turn (block i on XCD i % 8), so XCD x's blocks walk the x-th eighth
// of the list: the work of a few building boxes spreads over the whole chip,
// while neighbouring tiles, which read much the same data, share an XCD's
// L2.  (Over all slots, with the slot blocks of each XCD contiguous, a step
// in which a few small boxes built left most XCDs idle and launched a wave
// per 64 slots of the arena.)  The grid is at least 8 blocks, so every XCD
// share has a block.  f(s) is called for every slot of each tile; slots that
// do not build return at once.
template <class F>
__device__ __forceinline__ void for_building_tiles(const ArenaView &A, F f)
{
    constexpr int WPB = NT / 64;  // waves per block
    const int x = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int nbx = ((int)gridDim.x - x + 7) >> 3;  // blocks on this XCD
    const int T = A.plan_n[2];
    const int k0 = (int)(((long long)T * x) >> 3), k1 = (int)(((long long)T * (x + 1)) >> 3);
    const int lane = threadIdx.x & 63;
    for (int k = k0 + local * WPB + (int)(threadIdx.x >> 6); k < k1; k += nbx * WPB)
        f(A.tile_list[k] * 64 + lane);
}

__device__ __forceinline__ void keys_slot(const ArenaView &A, int s, int force_mask)
{
    if (s >= A.n) return;
    const int b = building_box(A, s, force_mask);
    if (b < 0) return;
    const BoxDesc &B = A.boxes[b];
    const double2 r = A.r[s];
    int cx = (int)(r.x * B.icwx), cy = (int)(r.y * B.icwy);
    cx = cx < 0 ? 0 : (cx >= B.ncx ? B.ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= B.ncy ? B.ncy - 1 : cy);
    A.keys_in[A.box_coff[b] + (s - B.part_off)] = (unsigned)(B.key_base + cy * B.ncx + cx);
}

__global__ __launch_bounds__(NT) void k_keys(ArenaView A, int force_mask)
{
    for_building_tiles(A, [&](int s) { keys_slot(A, s, force_mask); });
}

// Sorted compact position j (slot s of box b) takes the state of the slot at
// compact position vals[j] (same box: its key lies in the box's key range)
// into sorted scratch and records its cell.
__device__ __forceinline__ void permute_slot(const ArenaView &A, int s, int force_mask)
{
    if (s >= A.n) return;
    const int b = building_box(A, s, force_mask);
    if (b < 0) return;
    const BoxDesc &B = A.boxes[b];
    const int c0 = A.box_coff[b], j = c0 + (s - B.part_off);
    const int src = B.part_off + ((int)A.vals[j] - c0);
    A.r_s[s] = A.r[src];
    A.v_s[s] = A.v[src];
    A.vp_s[s] = A.vp[src];
    A.acc_s[s] = A.acc[src];
    A.u_s[s] = A.u[src];
    A.m_s[s] = A.m[src];
    A.up_s[s] = A.up[src];
    A.udot_s[s] = A.udot[src];
    A.id_s[s] = A.id[src];
    const int cell = (int)A.keys[j] - B.key_base;
    A.pcell[s] = (cell < 0 || cell >= B.ncx * B.ncy) ? 0 : cell;
}

// cell_start over the key range of each building box: cell_start[c] = first
// slot whose key >= c, and the end marker key_base + ncx*ncy = the slot past
// the box (the next segment's first slot: a building next box writes the same
// value there).
__device__ __forceinline__ void cell_start_slot(const ArenaView &A, int s, int force_mask)
{
    if (s >= A.n) return;
    const int b = building_box(A, s, force_mask);
    if (b < 0) return;
    const BoxDesc &B = A.boxes[b];
    const int i = s - B.part_off, j = A.box_coff[b] + i;
    const int key = (int)A.keys[j];
    const int kp = i == 0 ? B.key_base - 1 : (int)A.keys[j - 1];
    for (int c = kp + 1; c <= key; ++c) A.cell_start[c] = s;
    if (i == B.n - 1) {
        const int kend = B.key_base + B.ncx * B.ncy;
        for (int c = key + 1; c <= kend; ++c) A.cell_start[c] = s + 1;
    }
}

// The permutation, and with cells the cell starts from the sorted keys
// too (the radix-sort path; the per-box sort writes its own), in one
// launch: both walk the building slots and read only the sort's output.
__global__ __launch_bounds__(NT) void k_permute(ArenaView A, int force_mask, int cells)
{
    for_building_tiles(A, [&](int s) {
        if (cells) cell_start_slot(A, s, force_mask);
        permute_slot(A, s, force_mask);
    });
}

// Per-box counting sort for arenas of small boxes (every box with at most
// BS_MAX_CELLS cells and BS_MAX_N slots): one block per building box does
// what k_keys, the radix sort and the cell starts do for it, in LDS.  The order
// is the same stable one (by cell, then by compact position), so the two
// pipelines give identical lists.  Ranks come from wave ballots and per-wave
// counters (each wave owns one counter row and walks its slice of the box in
// order), so there are no atomics and the result is deterministic.  1024
// threads and 1024 cells (40 KB of LDS): the block of the largest box sets
// the kernel's time, and with 512 threads and 2048 cells it took twice as
// long per slot; boxes of up to 16,384 slots have at most 784 cells.
static constexpr int BS_THREADS = 1024;
static constexpr int BS_WAVES = BS_THREADS / 64;
static constexpr int BS_CELL_BITS = 10;
static constexpr int BS_MAX_CELLS = 1 << BS_CELL_BITS;

__device__ __forceinline__ uint64_t lanes_below()
{
    const unsigned lane = threadIdx.x & 63u;
    return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

// Lanes of the wave holding the same cell (valid lanes only); every lane of
// the wave calls it.
__device__ __forceinline__ uint64_t match_cell(int c, bool valid)
{
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < BS_CELL_BITS; ++bit) {
        const bool x = (c >> bit) & 1;
        const uint64_t bb = __ballot(x);
        m &= x ? bb : ~bb;
    }
    return m;
}

__device__ __forceinline__ int cell_of(double2 r, const BoxDesc &B)
{
    int cx = (int)(r.x * B.icwx), cy = (int)(r.y * B.icwy);
    cx = cx < 0 ? 0 : (cx >= B.ncx ? B.ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= B.ncy ? B.ncy - 1 : cy);
    return cy * B.ncx + cx;
}

__global__ __launch_bounds__(BS_THREADS) void k_box_sort(ArenaView A, int force_mask, int nbox)
{
    __shared__ unsigned short cnt[BS_WAVES][BS_MAX_CELLS];
    __shared__ int start[BS_MAX_CELLS + 1];
    __shared__ int part[BS_THREADS];
    const int b = blockIdx.x;
    if (b >= nbox || !building(A.dyn[b], force_mask)) return;
    const BoxDesc &B = A.boxes[b];
    const int n = B.n, C = B.ncx * B.ncy;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    for (int c = tid; c < C; c += BS_THREADS)
        for (int v = 0; v < BS_WAVES; ++v) cnt[v][c] = 0;
    __syncthreads();
    // wave w walks slots [s0, s1) of the box, 64 at a time, in order
    const int S = ((n + BS_WAVES - 1) / BS_WAVES + 63) & ~63;
    const int s0 = min(w * S, n), s1 = min(s0 + S, n);
    for (int g = s0; g < s1; g += 64) {
        const int i = g + lane;
        const bool valid = i < s1;
        const int c = valid ? cell_of(A.r[B.part_off + i], B) : 0;
        const uint64_t m = match_cell(c, valid);
        if (valid && (m & lanes_below()) == 0) cnt[w][c] += (unsigned short)__popcll(m);
    }
    __syncthreads();
    // start[c] = slots in cells below c; cnt[w][c] becomes the count of cell
    // c in the slices of waves below w
    const int per = (C + BS_THREADS - 1) / BS_THREADS;
    const int c0 = min(tid * per, C), c1 = min(c0 + per, C);
    int sum = 0;
    for (int c = c0; c < c1; ++c)
        for (int v = 0; v < BS_WAVES; ++v) sum += cnt[v][c];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < BS_THREADS; off <<= 1) {  // inclusive scan
        const int x = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += x;
        __syncthreads();
    }
    int run = part[tid] - sum;
    for (int c = c0; c < c1; ++c) {
        start[c] = run;
        int acc = 0;
        for (int v = 0; v < BS_WAVES; ++v) {
            const int x = cnt[v][c];
            cnt[v][c] = (unsigned short)acc;
            acc += x;
        }
        run += acc;
    }
    if (tid == 0) start[C] = n;
    __syncthreads();
    // stable scatter: compact positions of the sorted order
    const int coff = A.box_coff[b];
    for (int g = s0; g < s1; g += 64) {
        const int i = g + lane;
        const bool valid = i < s1;
        const int c = valid ? cell_of(A.r[B.part_off + i], B) : 0;
        const uint64_t m = match_cell(c, valid);
        const int rank = __popcll(m & lanes_below());
        if (valid) {
            const int pos = start[c] + cnt[w][c] + rank;
            A.vals[coff + pos] = (unsigned)(coff + i);
            A.keys[coff + pos] = (unsigned)(B.key_base + c);
        }
        if (valid && rank == 0) cnt[w][c] += (unsigned short)__popcll(m);
    }
    for (int c = tid; c <= C; c += BS_THREADS) A.cell_start[B.key_base + c] = B.part_off + start[c];
}

// Candidate scan of one building slot over the stencil cells of its build
// cell, in stencil order and slot order within a cell (the list order).
// SINGLE: a box with a single-cell axis, whose candidates need the minimum
// image.  Four candidate positions are loaded before any is tested.
// (Scanning a stencil row's consecutive cells as one slot range, in fewer
// and fuller rounds, measured slower: its range set-up cost more registers
// and instructions than the part-empty rounds it saved.)
//
// Entries collect in a ring of two G-entry cells per side (short entries
// from position 0 up, long ones from lcap - 1 down; G = 4 slots, or with L16
// G = 8 offsets of the 16-bit form): ring index sh ? ns % 2G : 4G-1 - nl % 2G,
// so a long cell sits in its half in list order.  A round appends at most
// four entries per side, so at most one cell per side completes in it, and
// the round's end stores the completed cells, one 16-byte store each, while
// the other half fills.  The partial cells are stored at the end, merged
// when they share a cell.  A list that would pass lcap entries sets lost
// and ends with no entries.  The ring is the lane's column of an LDS array
// (words buf[w * NT], 0-7 short, 8-15 long, 16 the spare that rejected
// candidates write), so an append is one indexed LDS write on one path
// for short and long entries, taken by every candidate.  (Stored from inside the candidate loop, as
// each cell completed, the store path ran for almost every candidate, some
// lane of the wave completing a cell at nearly every one: 20% more VALU
// instructions per round, and the C2 build stage took 122 us against 115.)
// Positions past a
// count hold stale words, which the walks never read (ListRounds maps them
// to the slot itself).
//
// 16-bit offsets: a stencil cell is a contiguous slot range, so one base
// serves all its candidates: the slot offset t - s when the cell's offsets
// all fit, else taken modulo n (the box's slot count) to the representative
// nearest 0 of the cell's middle; the cell's first and last offsets then
// decide whether all fit (else over), and a cell taken modulo n sets
// wrapped.  Boxes of up to 32,768 slots never wrap, and a box of n <= 65,536
// slots never overflows while its cells hold fewer than (65,536 - n) / 2
// slots each.  The candidate loop adds one constant per entry, as the
// 32-bit form stores the slot itself.
template <bool SINGLE, bool L16>
__device__ __forceinline__ void list_walk(const ArenaView &A, const BoxDesc &Bu, int s, double2 ri, int cell,
                                          unsigned *buf, int &ns, int &nl, bool &lost, bool &wrapped, bool &over)
{
    constexpr int G = L16 ? 8 : 4;  // entries per cell (one 16-byte store)
    const int lcap = A.lcap;
    const int nb = wave_uniform(Bu.n);
    const double rnl2 = wave_uniform(Bu.rnl2), hsk2 = wave_uniform(Bu.hsk2);
    const double Lx = wave_uniform(Bu.Lx), Ly = wave_uniform(Bu.Ly);
    const double hLx = wave_uniform(Bu.hLx), hLy = wave_uniform(Bu.hLy);
    const int ncx = wave_uniform(Bu.ncx), ncy = wave_uniform(Bu.ncy);
    const int sx = wave_uniform(Bu.sx), sy = wave_uniform(Bu.sy);
    const int *cs = A.cell_start + wave_uniform(Bu.key_base);
    const int cx = cell % ncx, cy = cell / ncx;
    int fs = 0, fl = 0;  // cells stored per side
    // the list address is formed at each store (not held across the loop)
    auto put = [&](int q, uint4 g) {
        if (L16) *(uint4 *)(list_base16(A, s) + (size_t)q * CELL16) = g;
        else *(uint4 *)(list_base(A, s) + (size_t)q * LIST_ROW) = g;
    };
    auto ring = [&](int w0) { return make_uint4(buf[w0 * NT], buf[(w0 + 1) * NT], buf[(w0 + 2) * NT], buf[(w0 + 3) * NT]); };
    // 16-bit base of the cell's candidates (0 in the 32-bit form)
    auto base16 = [&](int t0, int t1) {
        int ks = 0;
        if (L16 && t1 > t0) {
            const int dlo = t0 - s, dhi = t1 - 1 - s, dm = (dlo + dhi) >> 1;
            int k = 0;  // the plain offsets when they fit, else the nearest modulo n
            if (dlo < -32768 || dhi > 32767) k = 2 * dm > nb ? -nb : (2 * dm < -nb ? nb : 0);
            if (dlo + k < -32768 || dhi + k > 32767) over = true;
            if (k) wrapped = true;
            ks = k - s;
        }
        return ks;
    };
    // candidates t0 .. t1 - 1 at image offset (rx, ry) from r_i
    auto scan = [&](int t0, int t1, double rx, double ry, int ks) {
        for (int t = t0; t < t1; t += 4) {
            double2 rj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) rj[u] = at32(A.r_s, min(t + u, t1 - 1));
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double dx = rj[u].x - rx, dy = rj[u].y - ry;
                if (SINGLE && sx == 0) {
                    dx = dx > hLx ? dx - Lx : dx;
                    dx = dx < -hLx ? dx + Lx : dx;
                }
                if (SINGLE && sy == 0) {
                    dy = dy > hLy ? dy - Ly : dy;
                    dy = dy < -hLy ? dy + Ly : dy;
                }
                const double r2 = rsq_exact(dx, dy);
                // every candidate writes one ring entry, a rejected one to
                // the spare word 2G (row 4G), so the append takes no branch
                // (bitwise, so that the compiler selects instead of branching)
                const int take = (int)(t + u < t1) & (int)(r2 < rnl2) & (int)(t + u != s);
                const int sh = (int)(r2 < hsk2);
                const int rs = ns & (2 * G - 1), rl = (4 * G - 1) - (nl & (2 * G - 1));
                int row = sh ? rs : rl;
                row = take ? row : 4 * G;
                if (L16) ((unsigned short *)(buf + (row >> 1) * NT))[row & 1] = (unsigned short)(t + u + ks);
                else buf[row * NT] = (unsigned)(t + u);
                ns += take & sh;
                nl += take & (sh ^ 1);
            }
            // the cells completed in this round (at most one per side);
            // past lcap entries nothing more is stored
            lost = lost || ns + nl > lcap;
            if (!lost && ns / G > fs) {
                put(fs, ring(4 * (fs & 1)));
                ++fs;
            }
            if (!lost && nl / G > fl) {
                put(lcap / G - 1 - fl, ring(12 - 4 * (fl & 1)));
                ++fl;
            }
        }
    };
    for (int oy = -sy; oy <= sy; ++oy) {
        const int yr = cy + oy;
        const int yy = yr < 0 ? yr + ncy : (yr >= ncy ? yr - ncy : yr);
        const double ry = ri.y - (yr < 0 ? -Ly : (yr >= ncy ? Ly : 0.0));
        for (int ox = -sx; ox <= sx; ++ox) {
            const int xr = cx + ox;
            const int xx = xr < 0 ? xr + ncx : (xr >= ncx ? xr - ncx : xr);
            const double rx = ri.x - (xr < 0 ? -Lx : (xr >= ncx ? Lx : 0.0));
            const int c = yy * ncx + xx;
            const int t0 = cs[c], t1 = cs[c + 1];
            scan(t0, t1, rx, ry, base16(t0, t1));
        }
    }
    // a list that overflowed keeps no entries (its box is flagged
    // HEALTH_LISTCAP and its run ends)
    if (lost) {
        ns = nl = 0;
        return;
    }
    // partial cells: short positions [0, ns % G) of cell ns / G, long
    // positions [G - nl % G, G) of cell lcap / G - 1 - nl / G; with ns + nl
    // <= lcap they never overlap, and a shared cell is stored once, each
    // position taken from the side that owns it
    const int ps = ns % G, pl = nl % G;
    const int qs = fs, ql = lcap / G - 1 - fl;
    const uint4 ws = ring(4 * (fs & 1)), wl = ring(12 - 4 * (fl & 1));
    if (ps && pl && qs == ql) {
        const unsigned sw[4] = {ws.x, ws.y, ws.z, ws.w}, lw[4] = {wl.x, wl.y, wl.z, wl.w};
        unsigned w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (L16) w[i] = ((2 * i < ps ? sw[i] : lw[i]) & 0xffffu) | ((2 * i + 1 < ps ? sw[i] : lw[i]) & 0xffff0000u);
            else w[i] = i < ps ? sw[i] : lw[i];
        }
        put(qs, make_uint4(w[0], w[1], w[2], w[3]));
    } else {
        if (ps) put(qs, ws);
        if (pl) put(ql, wl);
    }
}

// Neighbour list of a building slot: every other slot of the 3x3 cell
// stencil with r^2 < (r_cut + skin)^2, short entries (r^2 < (h + skin)^2)
// from the front, the rest from the back.  The sorted state is copied back
// into the state arrays and the build position kept for the skin check.
// Returns the slot's box (-1: not building) and, in v2 / a2, its squared
// predicted speed and acceleration, which build_tail folds into the box's
// maxima for its next schedule.
//
// A wave covers one 64-slot list tile, so every tile is written by one wave.
// (Over compact positions, which are not tile-aligned, two waves, often on
// two XCDs with their own L2, wrote most tiles and the kernel ran 2.4x
// slower.)
__device__ __forceinline__ int build_slot(const ArenaView &A, int s, int force_mask, int fp32, unsigned *buf,
                                          double &v2, double &a2)
{
    if (s >= A.n) return -1;
    const int b = building_box(A, s, force_mask);
    if (b < 0) return -1;
    BoxDyn &D = A.dyn[b];
    const BoxDesc &B = A.boxes[b];
    const double2 ri = A.r_s[s];
    const double2 vpi = A.vp_s[s], ai = A.acc_s[s];
    const double mi = A.m_s[s];
    A.r[s] = ri;
    A.v[s] = A.v_s[s];
    A.vp[s] = vpi;
    A.acc[s] = ai;
    A.u[s] = A.u_s[s];
    A.m[s] = mi;
    A.up[s] = A.up_s[s];
    A.udot[s] = A.udot_s[s];
    A.id[s] = A.id_s[s];
    A.r_ref[s] = ri;
    const int cell = A.pcell[s];
    if (fp32) fix_rec(ri, vpi, B, A.rf[s], A.rlo[s]);
    int ns = 0, nl = 0;
    bool lost = false, wrapped = false, over = false;
    // Box constants wave-uniform (scalar registers); each stencil cell's
    // periodic image is applied to r_i once per cell, so the candidate loop
    // needs no minimum image (axes with a single cell keep it).
    per_box(b, [&](int bu) {
        const BoxDesc &Bu = A.boxes[bu];
        const bool single = wave_uniform(Bu.sx) == 0 || wave_uniform(Bu.sy) == 0;
        if (wave_uniform(Bu.build16)) {
            if (single) list_walk<true, true>(A, Bu, s, ri, cell, buf, ns, nl, lost, wrapped, over);
            else list_walk<false, true>(A, Bu, s, ri, cell, buf, ns, nl, lost, wrapped, over);
        } else {
            if (single) list_walk<true, false>(A, Bu, s, ri, cell, buf, ns, nl, lost, wrapped, over);
            else list_walk<false, false>(A, Bu, s, ri, cell, buf, ns, nl, lost, wrapped, over);
        }
    });
    if (lost) flag_health(A.dyn, b, HEALTH_LISTCAP);
    // an offset beyond 16 bits: the box holds this step (every pass, the
    // integration and the bookkeeping skip it) until the runtime has
    // switched it to 32-bit entries; its next step builds again
    if (over) D.relist = 1;
    if (s == B.part_off) D.list16 = B.build16;  // the form the walks read
    A.cnt[s] = make_int2(ns | (wrapped ? CNT_WRAP : 0), nl);
    v2 = vpi.x * vpi.x + vpi.y * vpi.y;
    a2 = ai.x * ai.x + ai.y * ai.y;
    return b;
}

// The build's largest squared speed and acceleration per box, as step_tail
// gathers the review displacement: a wave wholly inside one building box
// stores its maxima in wbuild (k_box_step folds the box's waves); a wave
// shared by several boxes folds each building box's lanes into that box's
// dyn.vmax2 / amax2 with one atomic each.  (A max_bits per lane put up to
// 2 x 64 atomics per wave on the box's two words, and while the running
// maxima were still low most lanes issued theirs: the in-run builds of a
// 262,144-slot box ran 10-200 us over the 76 us of one at a settled
// maximum.)  Every lane of the wave calls this.
__device__ __forceinline__ void build_tail(const ArenaView &A, int s, int b, double v2, double a2)
{
    const int lane = (int)(threadIdx.x & 63);
    const int bfirst = __shfl(b, 0);
    if (bfirst >= 0 && bfirst == __shfl(b, 63)) {  // wholly inside one box
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            v2 = fmax(v2, __shfl_xor(v2, off));
            a2 = fmax(a2, __shfl_xor(a2, off));
        }
        if (lane == 0) A.wbuild[s >> 6] = make_double2(v2, a2);
        return;
    }
    uint64_t pending = __ballot(b >= 0);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int bu = __shfl(b, leader);
        const bool mine = b == bu;
        double mv = mine ? v2 : 0.0, ma = mine ? a2 : 0.0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            mv = fmax(mv, __shfl_xor(mv, off));
            ma = fmax(ma, __shfl_xor(ma, off));
        }
        if (lane == leader) {
            max_bits(&A.dyn[bu].vmax2, mv);
            max_bits(&A.dyn[bu].amax2, ma);
        }
        pending &= ~__ballot(mine);
    }
}

// Over the building tiles (for_building_tiles).
__global__ __launch_bounds__(NT) void k_build_lists(ArenaView A, int force_mask, int fp32)
{
    // each lane's entry ring (two cells per side: eight words short, eight
    // long, and a spare word for rejected candidates), one column per lane:
    // word w of lane t at lbuf[w * NT + t], bank-conflict free
    __shared__ unsigned lbuf[17 * NT];
    for_building_tiles(A, [&](int s) {
        double v2 = 0.0, a2 = 0.0;
        const int b = build_slot(A, s, force_mask, fp32, lbuf + threadIdx.x, v2, a2);
        build_tail(A, s, b, v2, a2);
    });
}

// ---------------------------------------------------------------------------
// Box constants of the pair passes.  A wave whose active lanes all work on
// one box (every wave but those straddling a box boundary) makes them
// wave-uniform, so they sit in scalar registers for the whole list walk and
// the loop issues no loads but the neighbour records.
// ---------------------------------------------------------------------------
struct BoxK {
    double Lx, Ly, hLx, hLy;
    double cut2, h2, hl2, cutl, cuth;  // cutl = min(cut2, hl2), cuth = min(cut2, h2)
    double ih, ihl, qh, qhl, fh, fhl, munif, rin, iLx, iLy;
    double ih3, ihl3;  // 3/h, 3/h_l: 1 + 3q = fma(r, 3/h, 1)
    int terms;
};

__device__ __forceinline__ BoxK box_k(const BoxDesc &B, bool uniform)
{
    BoxK k;
    k.Lx = B.Lx; k.Ly = B.Ly; k.hLx = B.hLx; k.hLy = B.hLy;
    k.cut2 = B.cut2; k.h2 = B.h2; k.hl2 = B.hl2;
    k.cutl = B.cut2 < B.hl2 ? B.cut2 : B.hl2;
    k.cuth = B.cut2 < B.h2 ? B.cut2 : B.h2;
    k.ih = B.ih; k.ihl = B.ihl; k.qh = B.qh; k.qhl = B.qhl; k.fh = B.fh; k.fhl = B.fhl;
    k.munif = B.munif;
    k.ih3 = 3.0 * B.ih; k.ihl3 = 3.0 * B.ihl;
    k.rin = B.rin;
    k.iLx = 1.0 / B.Lx;
    k.iLy = 1.0 / B.Ly;
    k.terms = B.terms;
    if (uniform) {
        k.Lx = wave_uniform(k.Lx); k.Ly = wave_uniform(k.Ly); k.hLx = wave_uniform(k.hLx); k.hLy = wave_uniform(k.hLy);
        k.cut2 = wave_uniform(k.cut2); k.h2 = wave_uniform(k.h2); k.hl2 = wave_uniform(k.hl2);
        k.cutl = wave_uniform(k.cutl); k.cuth = wave_uniform(k.cuth);
        k.ih = wave_uniform(k.ih); k.ihl = wave_uniform(k.ihl); k.qh = wave_uniform(k.qh); k.qhl = wave_uniform(k.qhl);
        k.fh = wave_uniform(k.fh); k.fhl = wave_uniform(k.fhl); k.munif = wave_uniform(k.munif);
        k.rin = wave_uniform(k.rin); k.iLx = wave_uniform(k.iLx); k.iLy = wave_uniform(k.iLy);
        k.ih3 = wave_uniform(k.ih3); k.ihl3 = wave_uniform(k.ihl3);
        k.terms = wave_uniform(k.terms);
    }
    return k;
}

// ===========================================================================
// FP64 form
// ===========================================================================

// True when no active lane's particle lies within rin of a side: then no
// listed neighbour is across a side, the raw separation r_j - r_i is the
// minimum image, and the walks skip the image test (same bits: the test
// would leave every such separation alone).
__device__ __forceinline__ bool wave_interior(double2 r, const BoxK &K)
{
    const bool in = r.x >= K.rin && r.x <= K.Lx - K.rin && r.y >= K.rin && r.y <= K.Ly - K.rin;
    return __ballot(!in) == 0ull;
}

// Minimum image of a listed pair's raw separation.  A listed pair is closer
// than rin / 2, so the raw separation lies within rin / 2 of 0 or of +-L,
// where rint(dx / L) is exactly 0 or +-1 and dx - k L is one rounding of
// the same difference min_image forms: the same bits in 3 operations per
// axis instead of a compare and two selects per side.
__device__ __forceinline__ void min_image_listed(double &dx, double &dy, const BoxK &K)
{
    dx = fma(-K.Lx, rint(dx * K.iLx), dx);
    dy = fma(-K.Ly, rint(dy * K.iLy), dy);
}

// sqrt(x) for x >= 0 in a few FMAs: the hardware reciprocal square root,
// one Goldschmidt step for (sqrt x, 1 / (2 sqrt x)) and one Newton
// correction of the root, within about 1 ulp of the rounded root (pair
// separations only; x below 1e-280 returns 0).
__device__ __forceinline__ double sqrt_pair(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    const double d = fma(-g, g, x);
    g = fma(d, h, g);
    return x >= 1e-280 ? g : 0.0;
}

__device__ __forceinline__ double vdw_p(double rho, double T, const BoxDesc &B)
{
    return (rho * B.kB * T) / (1.0 - rho * B.b);  // properties.py:41
}

struct Dens64 {
    double rho, rhol, gx, gy, ev;
};

// P1 list walk.  COH: accumulate the cohesive terms (needs velocities);
// UNIM: all masses equal (no mass fetch).
template <int U64, bool COH, bool UNIM, bool WRAP, int LF>
__device__ __forceinline__ Dens64 density64_walk(const ArenaView &A, const BoxK &K, const ListFmt &L, int s)
{
    const double2 ri = A.r[s], vi = A.vp[s];
    const int2 cr = A.cnt[s];
    const int2 c = make_int2(cr.x & CNT_MASK, cr.y);
    // the sums leave out the kernel constants (and, with UNIM, the mass);
    // they scale the totals at the end
    Dens64 d;
    d.rho = d.rhol = d.gx = d.gy = d.ev = 0.0;
    // U64 entries per round: all their loads are issued before any use so
    // several neighbour fetches are in flight per wavefront; one load brings
    // a round's list entries.
    ListRounds<U64, false, true, LF> rounds(A, s, c.x, L);
    for (int k0 = 0; k0 < c.x; k0 += U64) {
        int t[U64];
        double2 rj[U64], vv[U64];
        double mj[U64];
        rounds.take(k0, t);
#pragma unroll
        for (int u = 0; u < U64; ++u) {
            rj[u] = at32(A.r, t[u]);
            if (!UNIM) mj[u] = at32(A.m, t[u]);
            if (COH) vv[u] = at32(A.vp, t[u]);
        }
        rounds.next(k0);
#pragma unroll
        for (int u = 0; u < U64; ++u) {
            double dx = rj[u].x - ri.x, dy = rj[u].y - ri.y;
            if (WRAP) min_image_listed(dx, dy, K);
            const double r2 = rsq_exact(dx, dy);
            if (k0 + u < c.x && r2 < K.cut2) {
                const double m = UNIM ? 1.0 : mj[u];
                const double rr = sqrt_pair(r2);
                if (r2 < K.hl2) {
                    const double w = fma(-rr, K.ihl, 1.0), w2 = w * w;
                    d.rhol = fma(m * fma(rr, K.ihl3, 1.0), w2 * w, d.rhol);
                    if (COH) {
                        const double g = m * w2;
                        d.gx = fma(g, dx, d.gx);
                        d.gy = fma(g, dy, d.gy);
                        d.ev = fma(g, (vv[u].x - vi.x) * dx + (vv[u].y - vi.y) * dy, d.ev);
                    }
                }
                if (r2 < K.h2) {
                    const double w = fma(-rr, K.ih, 1.0);
                    d.rho = fma(m * fma(rr, K.ih3, 1.0), w * w * w, d.rho);
                }
            }
        }
    }
    // long entries: row lcap-1-k holds the k-th, so all lanes of a wave
    // read the same row at step k (coalesced), whatever their counts
    ListRounds<U64, true, true, LF> lrounds(A, s, c.y, L);
    for (int k0 = 0; k0 < c.y; k0 += U64) {
        int t[U64];
        double2 rj[U64], vv[U64];
        double mj[U64];
        lrounds.take(k0, t);
#pragma unroll
        for (int u = 0; u < U64; ++u) {
            rj[u] = at32(A.r, t[u]);
            if (!UNIM) mj[u] = at32(A.m, t[u]);
            if (COH) vv[u] = at32(A.vp, t[u]);
        }
        lrounds.next(k0);
#pragma unroll
        for (int u = 0; u < U64; ++u) {
            double dx = rj[u].x - ri.x, dy = rj[u].y - ri.y;
            if (WRAP) min_image_listed(dx, dy, K);
            const double r2 = rsq_exact(dx, dy);
            if (k0 + u < c.y && r2 < K.cutl) {
                const double m = UNIM ? 1.0 : mj[u];
                const double rr = sqrt_pair(r2);
                const double w = fma(-rr, K.ihl, 1.0), w2 = w * w;
                d.rhol = fma(m * fma(rr, K.ihl3, 1.0), w2 * w, d.rhol);
                if (COH) {
                    const double g = m * w2;
                    d.gx = fma(g, dx, d.gx);
                    d.gy = fma(g, dy, d.gy);
                    d.ev = fma(g, (vv[u].x - vi.x) * dx + (vv[u].y - vi.y) * dy, d.ev);
                }
            }
        }
    }
    const double mu = UNIM ? K.munif : 1.0, cf = K.fhl * mu;
    d.rho = fma(K.qh * mu, d.rho, K.qh);
    d.rhol = fma(K.qhl * mu, d.rhol, K.qhl);
    d.gx *= cf;
    d.gy *= cf;
    d.ev *= cf;
    return d;
}

// P1: rho_i = W(0,h) + sum_j m_j W(r_ij, h)  (self term not times m_i,
// properties.py:77-78), rho_lr likewise with h_l; then T, p, pco.
//
// The cohesive part of the rates rides along.  Its pair factor is
// pco_i/rho_lr,i^2 + pco_j/rho_lr,j^2 with pco = -a rho_lr^2
// (f_properties.py:119), i.e. the constant -2a, so it needs only the
// positions, masses and velocities this pass walks anyway:
//   a_co,i  = -2a sum_j m_j dwdx_hl,ij
//   du_co,i = -a  sum_j m_j (v_j - v_i) . dwdx_hl,ij      (dv . f / 2)
// and the rate pass only walks the short entries (r < h + skin).
template <int U64>
__global__ __launch_bounds__(PASS_NT_MAX) void k_density64(ArenaView A, int state_mask)
{
    const int s = xcd_block() * (int)blockDim.x + threadIdx.x;
    if (s >= A.n) return;
    const int b = A.slot_box[s];
    if (b < 0) return;
    const int st = A.dyn[b].state;
    if (!state_in(st, state_mask) || A.dyn[b].relist) return;
    Dens64 d;
    per_box(b, [&](int bu) {
        const BoxK K = box_k(A.boxes[bu], true);
        const ListFmt L = list_fmt(A, bu, s);
        const bool coh = (K.terms & 2) != 0, unim = K.munif > 0.0, inner = wave_interior(A.r[s], K);
        with_form(L, [&](auto lf) {
            constexpr int F = decltype(lf)::value;
            if (inner) {
                if (coh) d = unim ? density64_walk<U64, true, true, false, F>(A, K, L, s) : density64_walk<U64, true, false, false, F>(A, K, L, s);
                else d = unim ? density64_walk<U64, false, true, false, F>(A, K, L, s) : density64_walk<U64, false, false, false, F>(A, K, L, s);
            } else {
                if (coh) d = unim ? density64_walk<U64, true, true, true, F>(A, K, L, s) : density64_walk<U64, true, false, true, F>(A, K, L, s);
                else d = unim ? density64_walk<U64, false, true, true, F>(A, K, L, s) : density64_walk<U64, false, false, true, F>(A, K, L, s);
            }
        });
    });
    const BoxDesc &B = A.boxes[b];
    const bool coh = (B.terms & 2u) != 0;
    const double rho = d.rho, rhol = d.rhol;
    const double mi = A.m[s];
    double ui = A.up[s];
    if (st == BOX_INIT) {  // u = kB T0 - a rho (properties.py:44-46,119)
        ui = B.kB * B.T0 - B.a * rho;
        A.u[s] = ui;
        A.up[s] = ui;
    }
    double T = (ui + B.a * rho) / B.kB;  // properties.py:48-49
    T = T < 0.0 ? 0.0 : T;               // spam_complete_force.py:152
    const double p = vdw_p(rho, T, B);
    A.rho[s] = make_double2(rho, rhol);
    A.T[s] = T;
    // the stress record (m in .w for the rate pass) only when the gradient
    // pass will not write it; q/rho^2 is read only with conduction, when the
    // gradient pass writes it
    if (!(B.terms & 4u)) {
        const double ps = (B.terms & 1u) ? p / (rho * rho) : 0.0;
        A.S[s] = make_double4(ps, 0.0, ps, mi);
    }
    A.vt[s] = make_double2(mi / rho, T);
    A.aco[s] = coh ? make_double2(-2.0 * B.a * d.gx, -2.0 * B.a * d.gy) : make_double2(0.0, 0.0);
    A.uco[s] = coh ? -B.a * d.ev : 0.0;
    if (!(rho * B.b < 1.0)) flag_health(A.dyn, b, HEALTH_EOS_POLE);
    if (!isfinite(rho) || !isfinite(rhol) || !isfinite(ui)) flag_health(A.dyn, b, HEALTH_NONFINITE);
}

struct Grad64 {
    double g00, g01, g10, g11, gt0, gt1;
};

template <int U64, bool WRAP, int LF>
__device__ __forceinline__ Grad64 grad64_walk(const ArenaView &A, const BoxK &K, const ListFmt &L, int s)
{
    const double2 ri = A.r[s], vi = A.vp[s];
    const double Ti = A.T[s];
    const int ns = A.cnt[s].x & CNT_MASK;
    Grad64 g;
    g.g00 = g.g01 = g.g10 = g.g11 = g.gt0 = g.gt1 = 0.0;
    ListRounds<U64, false, false, LF> rounds(A, s, ns, L);
    for (int k0 = 0; k0 < ns; k0 += U64) {
        int t[U64];
        double2 rj[U64], vv[U64], vt[U64];
        rounds.take(k0, t);
#pragma unroll
        for (int u = 0; u < U64; ++u) {
            rj[u] = at32(A.r, t[u]);
            vv[u] = at32(A.vp, t[u]);
            vt[u] = at32(A.vt, t[u]);
        }
        rounds.next(k0);
#pragma unroll
        for (int u = 0; u < U64; ++u) {
            double dx = rj[u].x - ri.x, dy = rj[u].y - ri.y;
            if (WRAP) min_image_listed(dx, dy, K);
            const double r2 = rsq_exact(dx, dy);
            if (k0 + u < ns && r2 < K.cuth) {
                const double rr = sqrt_pair(r2);
                const double q = rr * K.ih, w = 1.0 - q;
                const double F = K.fh * w * w;
                const double gx = -F * dx, gy = -F * dy;
                const double vj = vt[u].x;
                const double dvx = vv[u].x - vi.x, dvy = vv[u].y - vi.y;
                const double dT = vt[u].y - Ti;
                g.g00 += vj * dvx * gx;
                g.g01 += vj * dvx * gy;
                g.g10 += vj * dvy * gx;
                g.g11 += vj * dvy * gy;
                g.gt0 += vj * dT * gx;
                g.gt1 += vj * dT * gy;
            }
        }
    }
    return g;
}

// P2: grad v and grad T with grad_i W_ij = -dwdx_ij (dwdx from r_j - r_i);
// viscous stress Pi = -eta (G + G^T - tr G I) - zeta tr G I (2-D) and heat
// flux q = -k grad T.  diag != 0 also stores grad v and q for download.
// Neighbour record: position, velocity, (m/rho, T) from P1.
template <int U64>
__global__ __launch_bounds__(PASS_NT_MAX) void k_grad64(ArenaView A, int state_mask, int diag)
{
    const int s = xcd_block() * (int)blockDim.x + threadIdx.x;
    if (s >= A.n) return;
    const int b = A.slot_box[s];
    if (b < 0) return;
    if (!state_in(A.dyn[b].state, state_mask) || A.dyn[b].relist) return;
    const BoxDesc &B = A.boxes[b];
    if (!diag && !(B.terms & 12u)) return;
    Grad64 g;
    per_box(b, [&](int bu) {
        const BoxK K = box_k(A.boxes[bu], true);
        const ListFmt L = list_fmt(A, bu, s);
        const bool inner = wave_interior(A.r[s], K);
        with_form(L, [&](auto lf) {
            constexpr int F = decltype(lf)::value;
            g = inner ? grad64_walk<U64, false, F>(A, K, L, s) : grad64_walk<U64, true, F>(A, K, L, s);
        });
    });
    const double Ti = A.T[s];
    const double rho = A.rho[s].x, r2i = rho * rho;
    double pxx = 0, pxy = 0, pyy = 0;
    if (B.terms & 4u) {
        const double tr = g.g00 + g.g11;
        pxx = -B.eta * (2.0 * g.g00 - tr) - B.zeta * tr;
        pyy = -B.eta * (2.0 * g.g11 - tr) - B.zeta * tr;
        pxy = -B.eta * (g.g01 + g.g10);
        const double pr = (B.terms & 1u) ? vdw_p(rho, Ti, B) : 0.0;
        A.S[s] = make_double4((pr + pxx) / r2i, pxy / r2i, (pr + pyy) / r2i, A.m[s]);
    }
    double qx = 0, qy = 0;
    if (B.terms & 8u) {
        qx = -B.kcond * g.gt0;
        qy = -B.kcond * g.gt1;
        A.Q[s] = make_double2(qx / r2i, qy / r2i);
    }
    if (diag) {
        A.gradv[s] = make_double4(g.g00, g.g01, g.g10, g.g11);
        A.jq[s] = make_double2(qx, qy);
    }
}

struct Rate64 {
    double ax, ay, ud;
};

template <int U64, bool COND, bool WRAP, int LF>
__device__ __forceinline__ Rate64 rates64_walk(const ArenaView &A, const BoxK &K, const ListFmt &L, int s, int ns)
{
    const double2 ri = A.r[s], vi = A.vp[s];
    const double4 Si = A.S[s];
    const double2 Qi = COND ? A.Q[s] : make_double2(0.0, 0.0);
    Rate64 o;
    o.ax = o.ay = o.ud = 0.0;
    ListRounds<U64, false, false, LF> rounds(A, s, ns, L);
    for (int k0 = 0; k0 < ns; k0 += U64) {
        int t[U64];
        double2 rj[U64], vv[U64], Qj[U64];
        double4 Sj[U64];
        rounds.take(k0, t);
#pragma unroll
        for (int u = 0; u < U64; ++u) {
            rj[u] = at32(A.r, t[u]);
            vv[u] = at32(A.vp, t[u]);
            Sj[u] = at32(A.S, t[u]);
            if (COND) Qj[u] = at32(A.Q, t[u]);
        }
        rounds.next(k0);
#pragma unroll
        for (int u = 0; u < U64; ++u) {
            double dx = rj[u].x - ri.x, dy = rj[u].y - ri.y;
            if (WRAP) min_image_listed(dx, dy, K);
            const double r2 = rsq_exact(dx, dy);
            if (k0 + u < ns && r2 < K.cuth) {
                const double rr = sqrt_pair(r2);
                const double q = rr * K.ih, w = 1.0 - q;
                const double F = K.fh * w * w;
                const double wx = F * dx, wy = F * dy;
                const double axx = Si.x + Sj[u].x, axy = Si.y + Sj[u].y, ayy = Si.z + Sj[u].z;
                const double fx = axx * wx + axy * wy;
                const double fy = axy * wx + ayy * wy;
                const double cc = COND ? (Qi.x + Qj[u].x) * wx + (Qi.y + Qj[u].y) * wy : 0.0;
                const double e = 0.5 * ((vv[u].x - vi.x) * fx + (vv[u].y - vi.y) * fy);
                const double mj = Sj[u].w;
                o.ax += mj * fx;
                o.ay += mj * fy;
                o.ud += mj * (e + cc);
            }
        }
    }
    return o;
}

// ===========================================================================
// FP32-compensated form: FP32 neighbour records (48-bit box-relative fixed
// point positions), FP32 long-range pair arithmetic with Kahan-compensated
// sums, FP64 short-range terms, EOS, integration and energy.
// ===========================================================================

using true_t = std::integral_constant<bool, true>;
using false_t = std::integral_constant<bool, false>;

struct Kahan {
    float s, c;
    __device__ __forceinline__ Kahan() : s(0.f), c(0.f) {}
    __device__ __forceinline__ explicit Kahan(float x) : s(x), c(0.f) {}
    __device__ __forceinline__ void add(float x)
    {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

// The FP32C walks' accumulation scheme on a plain sequence, for tests: the
// terms rounded to FP32, summed in FP32 rounds of `unroll` terms, each round
// added to a Kahan total (out2[0]); beside it the plain FP32 running sum of
// the same terms (out2[1]).  One work-item does the sums.
__global__ __launch_bounds__(64) void k_fp32_accumulate(long long n, const double *__restrict__ x, int unroll,
                                                        double *__restrict__ out2)
{
    if (threadIdx.x != 0) return;
    Kahan k;
    float plain = 0.f;
    for (long long i = 0; i < n; i += unroll) {
        float part = 0.f;
        for (int u = 0; u < unroll && i + u < n; ++u) {
            const float t = (float)x[i + u];
            part += t;
            plain += t;
        }
        k.add(part);
    }
    out2[0] = k.value();
    out2[1] = (double)plain;
}

struct BoxF {
    double sclx, scly;    // side / 2^32: the unit of the high 32 bits
    double s48x, s48y;    // side / 2^48: the unit of the full 48 bits
    // FP64 constants of the short-range terms (short density, gradients,
    // rates), which take FP64 arithmetic on exact separations
    double h2d, cutsd, ihd, ih3d, fhd;
    float cut2, cutl;
    float ihl, qhl, fhl, munif;
    float ihl3;  // 3/h_l: 1 + 3q = fma(r, 3/h_l, 1)
    int terms;
};

__device__ __forceinline__ BoxF boxf(const BoxDesc &B, bool uniform)
{
    BoxF f;
    f.sclx = B.Lx / FIX32;
    f.scly = B.Ly / FIX32;
    f.s48x = B.Lx / FIX48;
    f.s48y = B.Ly / FIX48;
    f.h2d = B.h2;
    f.cutsd = B.R2list < B.h2 ? B.R2list : B.h2;
    f.ihd = B.ih;
    f.ih3d = 3.0 * B.ih;
    f.fhd = B.fh;
    f.munif = (float)B.munif;
    f.cut2 = (float)B.cut2;
    f.cutl = (float)(B.cut2 < B.hl2 ? B.cut2 : B.hl2);
    f.ihl = (float)B.ihl; f.qhl = (float)B.qhl; f.fhl = (float)B.fhl;
    f.ihl3 = (float)(3.0 * B.ihl);
    f.terms = B.terms;
    if (uniform) {
        f.sclx = wave_uniform(f.sclx); f.scly = wave_uniform(f.scly);
        f.s48x = wave_uniform(f.s48x); f.s48y = wave_uniform(f.s48y);
        f.h2d = wave_uniform(f.h2d); f.cutsd = wave_uniform(f.cutsd); f.ihd = wave_uniform(f.ihd);
        f.ih3d = wave_uniform(f.ih3d); f.fhd = wave_uniform(f.fhd);
        f.cut2 = wave_uniform(f.cut2); f.cutl = wave_uniform(f.cutl);
        f.ihl = wave_uniform(f.ihl); f.qhl = wave_uniform(f.qhl); f.fhl = wave_uniform(f.fhl);
        f.ihl3 = wave_uniform(f.ihl3); f.munif = wave_uniform(f.munif);
        f.terms = wave_uniform(f.terms);
    }
    return f;
}

// Minimum-image separation of record j from record i (fix_rec): the signed
// 32-bit difference of the high words (wrapping at the box side), with LO
// also the low 16 bits, which makes it the exact 48-bit difference of the
// two recorded positions; then one FP64 scale.
template <bool LO>
__device__ __forceinline__ void sep_fix(const float4 &ri, unsigned loi, const float4 &rj, unsigned loj, const BoxF &F,
                                        double &dx, double &dy)
{
    const int hx = (int)(__float_as_uint(rj.x) - __float_as_uint(ri.x));
    const int hy = (int)(__float_as_uint(rj.y) - __float_as_uint(ri.y));
    if constexpr (LO) {
        const long long ex = (long long)hx * 65536 + ((int)(loj & 0xffffu) - (int)(loi & 0xffffu));
        const long long ey = (long long)hy * 65536 + ((int)(loj >> 16) - (int)(loi >> 16));
        dx = (double)ex * F.s48x;
        dy = (double)ey * F.s48y;
    } else {
        dx = (double)hx * F.sclx;
        dy = (double)hy * F.scly;
    }
}

struct Dens32 {
    double rho;  // short-range density: FP64 terms on exact separations
    Kahan rhol, gx, gy, ev;
};

// The FP32 walks take pair distances from the hardware square root
// (v_sqrt_f32, about 1 ulp; the correctly rounded sequence costs ~14 more
// instructions per entry and moved no error against the FP64 oracle).

// P1 walk of the FP32C form.  The long-range terms (rho_lr and the
// cohesive rates, ~80% of the entries) are FP32: the terms of one unrolled
// round (U32 list entries) add in FP32 and each round's partial sum, times
// the kernel constant it leaves out, goes into a Kahan-compensated total.
// The short-range density (r < h: the density the EOS turns into T and p,
// where a relaxed liquid amplifies any error about tenfold) takes FP64 terms
// on the exact separations of the same records; its constants scale the
// FP64 total at the end.  A short entry gathers rf and rlo (20 bytes), a
// long entry rf alone (16), or rf and rlo in boxes wider than the side
// up to which the high words keep the long-range terms within FP32 rounding
// (LOLONG).  UNIM: all masses equal, none fetched.
template <int U32, bool COH, bool UNIM, bool LOLONG, int LF>
__device__ __forceinline__ void density32_walk(const ArenaView &A, const BoxF &F, const ListFmt &L, int s, Dens32 &d)
{
    // the kernel constants (and, with UNIM, the mass) scale each round's
    // partial sums instead of every term
    const float cqhl = UNIM ? F.qhl * F.munif : F.qhl;
    const float cfhl = UNIM ? F.fhl * F.munif : F.fhl;
    const float4 ri = A.rf[s];
    const unsigned loi = A.rlo[s];
    const int2 cr = A.cnt[s];
    const int2 c = make_int2(cr.x & CNT_MASK, cr.y);
    ListRounds<U32, false, true, LF> rounds(A, s, c.x, L);
    for (int k0 = 0; k0 < c.x; k0 += U32) {
        int t[U32];
        float4 rj[U32];
        unsigned lj[U32];
        float mm[U32];
        rounds.take(k0, t);
#pragma unroll
        for (int u = 0; u < U32; ++u) {
            rj[u] = at32(A.rf, t[u]);
            lj[u] = at32(A.rlo, t[u]);
            if (!UNIM) mm[u] = (float)at32(A.m, t[u]);
        }
        rounds.next(k0);
        float grhol = 0.f, ggx = 0.f, ggy = 0.f, gev = 0.f;
#pragma unroll
        for (int u = 0; u < U32; ++u) {
            double dxd, dyd;
            sep_fix<true>(ri, loi, rj[u], lj[u], F, dxd, dyd);
            const float dx = (float)dxd, dy = (float)dyd;
            const float r2 = dx * dx + dy * dy;
            const float mj = UNIM ? 1.f : mm[u];
            if (k0 + u < c.x && r2 < F.cutl) {
                const float rr = __builtin_amdgcn_sqrtf(r2);
                const float w = fmaf(-rr, F.ihl, 1.f), w2 = w * w;
                grhol = fmaf(mj * fmaf(rr, F.ihl3, 1.f), w2 * w, grhol);
                if (COH) {
                    const float g = mj * w2;
                    ggx = fmaf(g, dx, ggx);
                    ggy = fmaf(g, dy, ggy);
                    gev = fmaf(g, (rj[u].z - ri.z) * dx + (rj[u].w - ri.w) * dy, gev);
                }
            }
            const double r2d = rsq_exact(dxd, dyd);
            if (k0 + u < c.x && r2d < F.h2d) {
                const double rr = sqrt_pair(r2d), w = fma(-rr, F.ihd, 1.0);
                d.rho = fma((double)mj * fma(rr, F.ih3d, 1.0), w * w * w, d.rho);
            }
        }
        d.rhol.add(cqhl * grhol);
        if (COH) {
            d.gx.add(cfhl * ggx);
            d.gy.add(cfhl * ggy);
            d.ev.add(cfhl * gev);
        }
    }
    // long entries: row lcap-1-k holds the k-th, so all lanes of a wave
    // read the same row at step k (coalesced), whatever their counts
    ListRounds<U32, true, true, LF> lrounds(A, s, c.y, L);
    for (int k0 = 0; k0 < c.y; k0 += U32) {
        int t[U32];
        float4 rj[U32];
        unsigned lj[U32];
        float mm[U32];
        lrounds.take(k0, t);
#pragma unroll
        for (int u = 0; u < U32; ++u) {
            rj[u] = at32(A.rf, t[u]);
            if (LOLONG) lj[u] = at32(A.rlo, t[u]);
            if (!UNIM) mm[u] = (float)at32(A.m, t[u]);
        }
        lrounds.next(k0);
        float grhol = 0.f, ggx = 0.f, ggy = 0.f, gev = 0.f;
#pragma unroll
        for (int u = 0; u < U32; ++u) {
            double dxd, dyd;
            sep_fix<LOLONG>(ri, loi, rj[u], LOLONG ? lj[u] : 0u, F, dxd, dyd);
            const float dx = (float)dxd, dy = (float)dyd;
            const float r2 = dx * dx + dy * dy;
            const float mj = UNIM ? 1.f : mm[u];
            if (k0 + u < c.y && r2 < F.cutl) {
                const float rr = __builtin_amdgcn_sqrtf(r2);
                const float w = fmaf(-rr, F.ihl, 1.f), w2 = w * w;
                grhol = fmaf(mj * fmaf(rr, F.ihl3, 1.f), w2 * w, grhol);
                if (COH) {
                    const float g = mj * w2;
                    ggx = fmaf(g, dx, ggx);
                    ggy = fmaf(g, dy, ggy);
                    gev = fmaf(g, (rj[u].z - ri.z) * dx + (rj[u].w - ri.w) * dy, gev);
                }
            }
        }
        d.rhol.add(cqhl * grhol);
        if (COH) {
            d.gx.add(cfhl * ggx);
            d.gy.add(cfhl * ggy);
            d.ev.add(cfhl * gev);
        }
    }
}

// P1 of the FP32C form: rf carries position and predicted velocity; the
// cohesive rate part (see k_density64) is accumulated with the densities.
template <int U32>
__global__ __launch_bounds__(PASS_NT_MAX) void k_density32(ArenaView A, int state_mask)
{
    const int s = xcd_block() * (int)blockDim.x + threadIdx.x;
    if (s >= A.n) return;
    const int b = A.slot_box[s];
    if (b < 0) return;
    const int st = A.dyn[b].state;
    if (!state_in(st, state_mask) || A.dyn[b].relist) return;
    const BoxDesc &B = A.boxes[b];
    const bool coh = (B.terms & 2u) != 0;
    Dens32 d;
    per_box(b, [&](int bu) {
        const BoxDesc &Bu = A.boxes[bu];
        const BoxF F = boxf(Bu, true);
        const ListFmt L = list_fmt(A, bu, s);
        const bool lolong = wave_uniform(Bu.lolong) != 0;
        d.rho = 0.0;
        d.rhol = Kahan(F.qhl);
        auto walk = [&](auto coh_t, auto unim_t) {
            constexpr bool C = decltype(coh_t)::value, M = decltype(unim_t)::value;
            with_form(L, [&](auto lf) {
                constexpr int LF = decltype(lf)::value;
                if (lolong) density32_walk<U32, C, M, true, LF>(A, F, L, s, d);
                else density32_walk<U32, C, M, false, LF>(A, F, L, s, d);
            });
        };
        const bool unim = F.munif > 0.f;
        if (F.terms & 2) {
            if (unim) walk(true_t{}, true_t{});
            else walk(true_t{}, false_t{});
        } else {
            if (unim) walk(false_t{}, true_t{});
            else walk(false_t{}, false_t{});
        }
    });
    // the short-range sum leaves out q_h (and the common mass)
    const double rd = fma(B.qh * (B.munif > 0.0 ? B.munif : 1.0), d.rho, B.qh), rl = d.rhol.value();
    const double mi = A.m[s];
    double ui = A.up[s];
    if (st == BOX_INIT) {
        ui = B.kB * B.T0 - B.a * rd;
        A.u[s] = ui;
        A.up[s] = ui;
    }
    double T = (ui + B.a * rd) / B.kB;
    T = T < 0.0 ? 0.0 : T;
    const double p = vdw_p(rd, T, B);
    A.rho[s] = make_double2(rd, rl);
    A.T[s] = T;
    // records the gradient pass rewrites are written here only when it will
    // not run (stress without viscosity; q/rho^2 without viscosity and
    // conduction)
    if (!(B.terms & 4u)) {
        const double ps = (B.terms & 1u) ? p / (rd * rd) : 0.0;
        A.S32[s] = make_float4((float)ps, 0.f, (float)ps, (float)mi);
    }
    const unsigned lo = A.rlo[s];
    if (!(B.terms & 12u)) A.q32[s] = QRec32{lo, 0.f, 0.f, 0.f};
    A.g32[s] = GRec32{lo, (float)(mi / rd), T};
    A.aco[s] = coh ? make_double2(-2.0 * B.a * d.gx.value(), -2.0 * B.a * d.gy.value()) : make_double2(0.0, 0.0);
    A.uco[s] = coh ? -B.a * d.ev.value() : 0.0;
    if (!(rd * B.b < 1.0)) flag_health(A.dyn, b, HEALTH_EOS_POLE);
    if (!isfinite(rd) || !isfinite(rl) || !isfinite(ui)) flag_health(A.dyn, b, HEALTH_NONFINITE);
}

struct Grad32 {
    double g00, g01, g10, g11, gt0, gt1;
};

// P2 walk of the FP32C form: FP32 neighbour records, FP64 arithmetic on the
// exact separations.  On CDNA4 an FP64 FMA issues at the FP32 rate, so the
// short-range terms keep their precision at no VALU premium beyond the
// conversions; FP32 rounding of these terms (and of the temperature
// difference) left up to 4e-6 of force and heating error on relaxed liquid
// states, where the terms cancel to a small net rate.  The temperature comes
// from the FP64 T of the density pass.  Per neighbour two 16-byte loads:
// rf and the density pass's record (rlo, m/rho, T); as three records (rlo,
// m/rho with a float T, the FP64 T) the pass was no faster than the FP64
// form's, every gather costing the same texture-path time whatever its
// width.
template <int U32, int LF>
__device__ __forceinline__ void grad32_walk(const ArenaView &A, const BoxF &F, const ListFmt &L, int s, Grad32 &g)
{
    const float4 ri = A.rf[s];
    const unsigned loi = A.rlo[s];
    const double vix = ri.z, viy = ri.w;
    const double Ti = A.T[s];
    const int ns = A.cnt[s].x & CNT_MASK;
    ListRounds<U32, false, false, LF> rounds(A, s, ns, L);
    for (int k0 = 0; k0 < ns; k0 += U32) {
        int t[U32];
        float4 rj[U32];
        GRec32 gj[U32];
        rounds.take(k0, t);
#pragma unroll
        for (int u = 0; u < U32; ++u) {
            rj[u] = at32(A.rf, t[u]);
            gj[u] = at32(A.g32, t[u]);
        }
        rounds.next(k0);
#pragma unroll
        for (int u = 0; u < U32; ++u) {
            double dx, dy;
            sep_fix<true>(ri, loi, rj[u], gj[u].lo, F, dx, dy);
            const double r2 = rsq_exact(dx, dy);
            if (k0 + u < ns && r2 < F.cutsd) {
                const double rr = sqrt_pair(r2), w = fma(-rr, F.ihd, 1.0);
                const double Fk = F.fhd * w * w;
                const double gx = -Fk * dx, gy = -Fk * dy;
                const double vj = gj[u].v;
                const double dvx = (double)rj[u].z - vix, dvy = (double)rj[u].w - viy;
                const double dT = gj[u].T - Ti;
                g.g00 += vj * dvx * gx;
                g.g01 += vj * dvx * gy;
                g.g10 += vj * dvy * gx;
                g.g11 += vj * dvy * gy;
                g.gt0 += vj * dT * gx;
                g.gt1 += vj * dT * gy;
            }
        }
    }
}

// P2 of the FP32C form.  Writes the stress record (S32, with viscosity) and
// the rate pass's q/rho^2 record.
template <int U32>
__global__ __launch_bounds__(PASS_NT_MAX) void k_grad32(ArenaView A, int state_mask, int diag)
{
    const int s = xcd_block() * (int)blockDim.x + threadIdx.x;
    if (s >= A.n) return;
    const int b = A.slot_box[s];
    if (b < 0) return;
    if (!state_in(A.dyn[b].state, state_mask) || A.dyn[b].relist) return;
    const BoxDesc &B = A.boxes[b];
    if (!diag && !(B.terms & 12u)) return;
    Grad32 g = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    per_box(b, [&](int bu) {
        const BoxF F = boxf(A.boxes[bu], true);
        const ListFmt L = list_fmt(A, bu, s);
        with_form(L, [&](auto lf) { grad32_walk<U32, decltype(lf)::value>(A, F, L, s, g); });
    });
    const double G00 = g.g00, G01 = g.g01, G10 = g.g10, G11 = g.g11;
    const double rho = A.rho[s].x, r2i = rho * rho;
    double qx = 0, qy = 0;
    if (B.terms & 4u) {
        const double tr = G00 + G11;
        const double pxx = -B.eta * (2.0 * G00 - tr) - B.zeta * tr;
        const double pyy = -B.eta * (2.0 * G11 - tr) - B.zeta * tr;
        const double pxy = -B.eta * (G01 + G10);
        const double pr = (B.terms & 1u) ? vdw_p(rho, A.T[s], B) : 0.0;
        A.S32[s] = make_float4((float)((pr + pxx) / r2i), (float)(pxy / r2i), (float)((pr + pyy) / r2i),
                               (float)A.m[s]);
    }
    if (B.terms & 8u) {
        qx = -B.kcond * g.gt0;
        qy = -B.kcond * g.gt1;
    }
    A.q32[s] = QRec32{A.rlo[s], (float)(qx / r2i), (float)(qy / r2i), 0.f};
    if (diag) {
        A.gradv[s] = make_double4(G00, G01, G10, G11);
        A.jq[s] = make_double2(qx, qy);
    }
}

struct Rate32 {
    double ax, ay, ud;
};

// P3 walk of the FP32C form: FP32 neighbour records, FP64 arithmetic on the
// exact separations (see grad32_walk).  Per neighbour: rf, (S, m) and rlo,
// the latter, with conduction, in one record with q/rho^2.
template <int U32, bool COND, int LF>
__device__ __forceinline__ void rates32_walk(const ArenaView &A, const BoxF &F, const ListFmt &L, int s, int ns,
                                             Rate32 &o)
{
    const float4 ri = A.rf[s];
    const unsigned loi = A.rlo[s];
    float2 qi = make_float2(0.f, 0.f);
    if (COND) {
        const QRec32 q = A.q32[s];
        qi = make_float2(q.qx, q.qy);
    }
    const float4 Si = A.S32[s];
    ListRounds<U32, false, false, LF> rounds(A, s, ns, L);
    for (int k0 = 0; k0 < ns; k0 += U32) {
        int t[U32];
        float4 rj[U32], Sj[U32];
        unsigned lj[U32];
        float2 qj[U32];
        rounds.take(k0, t);
#pragma unroll
        for (int u = 0; u < U32; ++u) {
            rj[u] = at32(A.rf, t[u]);
            Sj[u] = at32(A.S32, t[u]);
            if (COND) {  // the low position bits ride in the conduction record
                const QRec32 q = at32(A.q32, t[u]);
                lj[u] = q.lo;
                qj[u] = make_float2(q.qx, q.qy);
            } else {
                lj[u] = at32(A.rlo, t[u]);
                qj[u] = make_float2(0.f, 0.f);
            }
        }
        rounds.next(k0);
#pragma unroll
        for (int u = 0; u < U32; ++u) {
            double dx, dy;
            sep_fix<true>(ri, loi, rj[u], lj[u], F, dx, dy);
            const double r2 = rsq_exact(dx, dy);
            if (k0 + u < ns && r2 < F.cutsd) {
                const double rr = sqrt_pair(r2), w = fma(-rr, F.ihd, 1.0);
                const double Fk = F.fhd * w * w;
                const double wx = Fk * dx, wy = Fk * dy;
                const double axx = (double)Si.x + (double)Sj[u].x, axy = (double)Si.y + (double)Sj[u].y;
                const double ayy = (double)Si.z + (double)Sj[u].z;
                const double fx = axx * wx + axy * wy;
                const double fy = axy * wx + ayy * wy;
                const double cc = COND ? ((double)qi.x + (double)qj[u].x) * wx + ((double)qi.y + (double)qj[u].y) * wy
                                       : 0.0;
                const double mj = Sj[u].w;
                const double e = 0.5 * (((double)rj[u].z - (double)ri.z) * fx + ((double)rj[u].w - (double)ri.w) * fy);
                o.ax += mj * fx;
                o.ay += mj * fy;
                o.ud += mj * (e + cc);
            }
        }
    }
}

// ===========================================================================
// P3 rates with the P4 leapfrog fused, shared by both forms (FP64 state)
// ===========================================================================

// Running boxes kick by dt/2 at k = 0 and k = nsteps and by dt otherwise,
// and drift (with periodic wrap) while k < nsteps.  Between steps v and u are
// half-step values; vp / up get the prediction to the next integer time.
// The energy sum m (v^2/2 + u) of each slot is captured at the first and
// last synchronised state.
//
// The kick and drift run in the rate pass, right after a slot's rates,
// while other slots of the same launch still gather this slot's position
// and predicted velocity (and FP32C record): those go to the second state
// buffers (r_nx, vp_nx, rf_nx, rlo_nx), which the runtime makes the current
// ones after the step; v, u and up are read by their own slot only and are
// updated in place.  A live slot of a box that takes no step in this launch
// (initialising, done, held) is carried over into the second buffers.
//
// Leapfrog step of slot s of running box b with this evaluation's rates
// (a, ud); returns the squared displacement since the box's last build (-1
// at the closing half kick, which moves nothing).  On the box's first and
// last step, cap is set and e holds the slot's energy.
__device__ __forceinline__ double integrate_slot(const ArenaView &A, int fp32, int s, int b, double2 a, double ud,
                                                 double &e, bool &cap)
{
    const BoxDyn &D = A.dyn[b];
    const BoxDesc &B = A.boxes[b];
    const int64_t k = D.k, K = D.nsteps;
    const double kick = (k == 0 || k == K) ? 0.5 * B.dt : B.dt;
    const double2 v = A.v[s];
    const double u = A.u[s], m = A.m[s];
    if (k == 0) {
        e = m * (0.5 * (v.x * v.x + v.y * v.y) + u);
        A.e[s] = e;
        cap = true;
    }
    const double vx = v.x + kick * a.x, vy = v.y + kick * a.y;
    const double un = u + kick * ud;
    A.v[s] = make_double2(vx, vy);
    A.u[s] = un;
    const double2 r = A.r[s];
    if (k == K) {
        e = m * (0.5 * (vx * vx + vy * vy) + un);
        A.e[s] = e;
        cap = true;
        A.r_nx[s] = r;
        A.vp_nx[s] = make_double2(vx, vy);
        A.up[s] = un;
        if (fp32) fix_rec(r, make_double2(vx, vy), B, A.rf_nx[s], A.rlo_nx[s]);
        return -1.0;
    }
    double2 rn;
    rn.x = wrap_coord(r.x + B.dt * vx, B.Lx);
    rn.y = wrap_coord(r.y + B.dt * vy, B.Ly);
    A.r_nx[s] = rn;
    // a non-finite velocity (or energy) of a particle without neighbours
    // reaches no rate; catch it in the state
    if (!isfinite(rn.x) || !isfinite(rn.y) || !isfinite(un)) flag_health(A.dyn, b, HEALTH_NONFINITE);
    const double hdt = 0.5 * B.dt;  // predictor to the next integer time
    const double2 vp = make_double2(vx + hdt * a.x, vy + hdt * a.y);
    A.vp_nx[s] = vp;
    A.up[s] = un + hdt * ud;
    const double2 r0 = A.r_ref[s];
    double ddx = rn.x - r0.x, ddy = rn.y - r0.y;
    min_image(ddx, ddy, B);
    const double d2 = ddx * ddx + ddy * ddy;
    if (d2 > B.hskin2) flag_health(A.dyn, b, HEALTH_SKIN);
    if (fp32) fix_rec(rn, vp, B, A.rf_nx[s], A.rlo_nx[s]);
    return d2;
}

// A live slot whose box takes no step in this launch: its state, unchanged,
// into the second buffers.
__device__ __forceinline__ void carry_slot(const ArenaView &A, int fp32, int s)
{
    A.r_nx[s] = A.r[s];
    A.vp_nx[s] = A.vp[s];
    if (fp32) {
        A.rf_nx[s] = A.rf[s];
        A.rlo_nx[s] = A.rlo[s];
    }
}

// A running box re-examines its list at the step before a scheduled build
// (k_box_step, extend_steps); on that step the step pass gathers the largest
// squared displacement of the box's slots: per wave in wdisp for the waves
// wholly inside the box, in its dyn.dmax2 for the waves it shares.
__device__ __forceinline__ bool list_review(const BoxDyn &D)
{
    // not at the closing half kick (k = nsteps), which moves nothing; the
    // step before it reviews as any other, so the schedule of a run does
    // not depend on where its closing step lies (an open run has none)
    return D.state == BOX_RUN && !building(D, 0) && D.k < D.nsteps && D.k + 1 >= D.next_rebuild;
}

// The step pass's tail for slot s of box b (-1: no box): the leapfrog step
// (on: the box runs and took its rates in this launch) or the carry-over,
// then, on a box's review step, the largest squared displacement of its
// lanes.  A wave holds 64 consecutive slots and a box a contiguous slot
// range, so a wave whose first and last lanes share a box lies wholly inside
// it: it stores its maximum in wdisp.  A wave shared by several boxes (at
// most two per box) folds each reviewing box's lanes into that box's
// dyn.dmax2 with one atomic.  Each box thus sees only its own slots, so its
// schedule, and with it its results, do not depend on the boxes beside it.
// (One atomic per wave into a box field made every review step of a large
// box slow, the waves queueing on one address.)  Every lane of the wave
// calls this (no early return before the shuffles).
__device__ __forceinline__ void step_tail(const ArenaView &A, int fp32, int s, int b, bool on, double2 a, double ud)
{
    double d2 = 0.0, e = 0.0;
    bool review = false, cap = false;
    if (s < A.n && b >= 0) {
        if (on) {
            d2 = fmax(integrate_slot(A, fp32, s, b, a, ud, e, cap), 0.0);
            review = list_review(A.dyn[b]);
        } else {
            carry_slot(A, fp32, s);
        }
    }
    const int lane = (int)(threadIdx.x & 63);
    const int bfirst = __shfl(b, 0);
    const bool whole = bfirst >= 0 && bfirst == __shfl(b, 63);  // wholly inside one box
    // a box's first or last step: a whole wave's energy sum (fixed shuffle
    // order, every lane the same total) for k_box_step; the box's lanes in
    // shared waves it sums from e itself
    if (whole && __ballot(cap) != 0ull) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) e += __shfl_xor(e, off);
        if (lane == 0) A.wenergy[s >> 6] = e;
    }
    uint64_t pending = __ballot(review);
    if (pending == 0ull) return;
    if (whole) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) d2 = fmax(d2, __shfl_xor(d2, off));
        if (lane == 0) A.wdisp[s >> 6] = d2;
        return;
    }
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int bu = __shfl(b, leader);
        const bool mine = review && b == bu;
        double m = mine ? d2 : 0.0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
        if (lane == leader) max_bits(&A.dyn[bu].dmax2, m);
        pending &= ~__ballot(mine);
    }
}

// P3: rates over the short entries, plus the cohesive part from P1.
//   f   = (S_i + S_j) . dwdx_h           (S = (p I + Pi) / rho^2)
//   a_i = sum_j m_j f + a_co,i;   udot_i = sum_j m_j (dv . f / 2 + (Q_i + Q_j) . dwdx_h) + du_co,i
// Neighbour record: position, velocity, (Sxx, Sxy, Syy, m), Q.  step: the
// stepping launch, which then takes the leapfrog step (step_tail); the
// stand-alone evaluations only store the rates.
template <int U64>
__global__ __launch_bounds__(PASS_NT_MAX) void k_rates64(ArenaView A, int state_mask, int step)
{
    const int s = xcd_block() * (int)blockDim.x + threadIdx.x;
    int b = -1;
    bool on = false;
    if (s < A.n) {
        b = A.slot_box[s];
        on = b >= 0 && state_in(A.dyn[b].state, state_mask) && !A.dyn[b].relist;
    }
    double2 a = make_double2(0.0, 0.0);
    double ud = 0.0;
    if (on) {
        const BoxDesc &B = A.boxes[b];
        const int ns = (B.terms & 13u) ? (A.cnt[s].x & CNT_MASK) : 0;
        Rate64 o;
        per_box(b, [&](int bu) {
            const BoxK K = box_k(A.boxes[bu], true);
            const ListFmt L = list_fmt(A, bu, s);
            const bool inner = wave_interior(A.r[s], K), cond = (K.terms & 8) != 0;
            with_form(L, [&](auto lf) {
                constexpr int F = decltype(lf)::value;
                if (inner)
                    o = cond ? rates64_walk<U64, true, false, F>(A, K, L, s, ns) : rates64_walk<U64, false, false, F>(A, K, L, s, ns);
                else
                    o = cond ? rates64_walk<U64, true, true, F>(A, K, L, s, ns) : rates64_walk<U64, false, true, F>(A, K, L, s, ns);
            });
        });
        const double2 ac = A.aco[s];
        a = make_double2(o.ax + ac.x, o.ay + ac.y);
        ud = o.ud + A.uco[s];
        if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(ud)) flag_health(A.dyn, b, HEALTH_NONFINITE);
        A.acc[s] = a;
        A.udot[s] = ud;
    }
    if (step) step_tail(A, 0, s, b, on, a, ud);
}

// P3 of the FP32C form over the short entries; adds the cohesive part from
// P1; step as in k_rates64.
template <int U32>
__global__ __launch_bounds__(PASS_NT_MAX) void k_rates32(ArenaView A, int state_mask, int step)
{
    const int s = xcd_block() * (int)blockDim.x + threadIdx.x;
    int b = -1;
    bool on = false;
    if (s < A.n) {
        b = A.slot_box[s];
        on = b >= 0 && state_in(A.dyn[b].state, state_mask) && !A.dyn[b].relist;
    }
    double2 a = make_double2(0.0, 0.0);
    double ud = 0.0;
    if (on) {
        const BoxDesc &B = A.boxes[b];
        const int ns = (B.terms & 13u) ? (A.cnt[s].x & CNT_MASK) : 0;
        Rate32 o = {0.0, 0.0, 0.0};
        per_box(b, [&](int bu) {
            const BoxF F = boxf(A.boxes[bu], true);
            const ListFmt L = list_fmt(A, bu, s);
            with_form(L, [&](auto lf) {
                constexpr int LF = decltype(lf)::value;
                if (F.terms & 8) rates32_walk<U32, true, LF>(A, F, L, s, ns, o);
                else rates32_walk<U32, false, LF>(A, F, L, s, ns, o);
            });
        });
        const double2 ac = A.aco[s];
        a = make_double2(o.ax + ac.x, o.ay + ac.y);
        ud = o.ud + A.uco[s];
        if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(ud)) flag_health(A.dyn, b, HEALTH_NONFINITE);
        A.acc[s] = a;
        A.udot[s] = ud;
    }
    if (step) step_tail(A, 1, s, b, on, a, ud);
}

// The step pass as a launch of its own (the schedule's unfused form): the
// same leapfrog step and carry-over as step_tail in the rate pass, from the
// rates that pass stored.  The same bits either way.
__global__ __launch_bounds__(NT) void k_leapfrog(ArenaView A, int fp32)
{
    const int s = xcd_block() * NT + threadIdx.x;
    int b = -1;
    bool on = false;
    double2 a = make_double2(0.0, 0.0);
    double ud = 0.0;
    if (s < A.n) {
        b = A.slot_box[s];
        on = b >= 0 && A.dyn[b].state == BOX_RUN && !A.dyn[b].relist;
        if (on) {
            a = A.acc[s];
            ud = A.udot[s];
        }
    }
    step_tail(A, fp32, s, b, on, a, ud);
}

// ===========================================================================
// per-box bookkeeping
// ===========================================================================

// Steps a box may reuse the list it just built: the largest c with
//   c dt vmax + (c dt)^2 amax <= 0.4 skin    (twice the kinematic bound on
// the acceleration term, 80 % of the half skin), plus the build step; vmax,
// amax the largest speed and acceleration at the build.
__device__ __forceinline__ int64_t reuse_steps(const BoxDesc &B, double vmax, double amax)
{
    const double C = 0.4 * B.skin, Bq = B.dt * vmax, Aq = B.dt * B.dt * amax;
    double c;
    if (Aq > 0.0) c = (-Bq + sqrt(Bq * Bq + 4.0 * Aq * C)) / (2.0 * Aq);
    else if (Bq > 0.0) c = C / Bq;
    else c = 1e9;
    int64_t K = (int64_t)floor(c) + 1;
    if (amax == 0.0 && K > 2) K = 2;  // no rates yet: re-measure soon
    return K < 1 ? 1 : (K > 64 ? 64 : K);
}

// Steps a box may still reuse its list at a scheduled build step: the
// largest c with  c dt v + (c dt)^2 a <= 0.8 (skin/2 - d)  for the largest
// current displacement d from the build positions, the acceleration a
// measured at the last build and the speed v that bounds the particles now:
// the speed measured at the last build plus what that acceleration adds over
// the list's age (reuse_steps' bound applied to the margin that is left),
// capped so that a list lives at most 64 steps; 0 = build now.  Exceeding
// half the skin anyway raises HEALTH_SKIN in the step pass (integrate_slot).
__device__ __forceinline__ int64_t extend_steps(const BoxDesc &B, const BoxDyn &D, double dmax2)
{
    // a list built before any rates existed (acceleration 0, as reuse_steps
    // notes) is never extended: its bound has no acceleration term
    if (!(D.alast > 0.0)) return 0;
    const double d = sqrt(dmax2);
    const double C = 0.8 * (0.5 * B.skin - d);
    if (!(C > 0.0)) return 0;
    const int64_t age = D.k + 1 - D.built_at;  // list age at the next step
    const double Bq = B.dt * (D.vlast + (double)age * B.dt * D.alast), Aq = B.dt * B.dt * D.alast;
    const double c = (-Bq + sqrt(Bq * Bq + 4.0 * Aq * C)) / (2.0 * Aq);
    const int64_t room = 64 - age;
    int64_t e = (int64_t)floor(c);
    e = e > room ? room : e;
    return e < 0 ? 0 : e;
}

// Per-box step bookkeeping, one block per box.  A running box at k = 0 or
// k = nsteps first sums its slots' energies m (v^2/2 + u) into E0 / E1
// (fixed order: four strided partial sums per thread, then a tree).  Then
// INIT -> RUN (k = 0); RUN: k++ until k == nsteps, then DONE.  Boxes that
// built this step get their next build step from their own speed and
// acceleration, so a box's schedule never depends on the other boxes.
// With dyn_host set, thread 0 leaves the box's state there last: after a
// step the host reads it without a device-to-host copy.
static constexpr int NT_BOX = 1024;

__global__ __launch_bounds__(NT_BOX) void k_box_step(ArenaView A, int nbox)
{
    const int b = blockIdx.x;
    if (b >= nbox) return;
    BoxDyn &D = A.dyn[b];
    if (D.relist) {  // held: no step taken, the next one builds again
        if (threadIdx.x == 0 && A.dyn_host) A.dyn_host[b] = D;
        return;
    }
    const int state = D.state;
    const int64_t k = D.k, K = D.nsteps;
    const bool capture = state == BOX_RUN && (k == 0 || k == K);
    const bool review = list_review(D);
    const bool built = building(D, 0);
    __shared__ double red[NT_BOX];
    double dmax2 = 0.0, vmax2 = 0.0, amax2 = 0.0;
    if (built) {  // the build's largest squared speed and acceleration: the box's whole waves, then its shared ones
        const BoxDesc &B = A.boxes[b];
        const int w0 = (B.part_off + 63) >> 6, w1 = ((B.part_off + B.n) >> 6) - 1;
        double mv = 0.0, ma = 0.0;
        for (int w = w0 + (int)threadIdx.x; w <= w1; w += NT_BOX) {
            const double2 x = A.wbuild[w];
            mv = fmax(mv, x.x);
            ma = fmax(ma, x.y);
        }
        red[threadIdx.x] = mv;
        __syncthreads();
        for (int off = NT_BOX / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
            __syncthreads();
        }
        vmax2 = fmax(red[0], __longlong_as_double((long long)D.vmax2));
        __syncthreads();
        red[threadIdx.x] = ma;
        __syncthreads();
        for (int off = NT_BOX / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
            __syncthreads();
        }
        amax2 = fmax(red[0], __longlong_as_double((long long)D.amax2));
        __syncthreads();  // red is reused below
    }
    if (review) {  // largest squared displacement: the box's whole waves, then its shared ones
        const BoxDesc &B = A.boxes[b];
        const int w0 = (B.part_off + 63) >> 6, w1 = ((B.part_off + B.n) >> 6) - 1;
        double m = 0.0;
        for (int w = w0 + (int)threadIdx.x; w <= w1; w += NT_BOX) m = fmax(m, A.wdisp[w]);
        red[threadIdx.x] = m;
        __syncthreads();
        for (int off = NT_BOX / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
            __syncthreads();
        }
        dmax2 = fmax(red[0], __longlong_as_double((long long)D.dmax2));
        __syncthreads();  // red is reused below
    }
    if (capture) {
        // the step pass's sums of the box's whole waves (wenergy), then the
        // box's slots in waves it shares (at most 63 at each end)
        const BoxDesc &B = A.boxes[b];
        const int w0 = (B.part_off + 63) >> 6, w1 = ((B.part_off + B.n) >> 6) - 1;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int w = w0 + (int)threadIdx.x;
        for (; w + 3 * NT_BOX <= w1; w += 4 * NT_BOX) {
            a0 += A.wenergy[w];
            a1 += A.wenergy[w + NT_BOX];
            a2 += A.wenergy[w + 2 * NT_BOX];
            a3 += A.wenergy[w + 3 * NT_BOX];
        }
        for (; w <= w1; w += NT_BOX) a0 += A.wenergy[w];
        red[threadIdx.x] = (a0 + a1) + (a2 + a3);
        __syncthreads();
        for (int off = NT_BOX / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
    }
    __syncthreads();  // every thread has read D before thread 0 updates it
    if (threadIdx.x != 0) return;
    if (capture) {
        const BoxDesc &B = A.boxes[b];
        const int s0 = B.part_off, s1 = B.part_off + B.n;
        const int h1 = min(((s0 + 63) >> 6) << 6, s1), t0 = max((s1 >> 6) << 6, h1);
        double ends = 0.0;
        for (int s = s0; s < h1; ++s) ends += A.e[s];
        for (int s = t0; s < s1; ++s) ends += A.e[s];
        const double tot = red[0] + ends;
        if (k == 0) D.E0 = tot;
        else D.E1 = tot;
    }
    if (built) {
        D.vlast = sqrt(vmax2);
        D.alast = sqrt(amax2);
        const int64_t nK = reuse_steps(A.boxes[b], D.vlast, D.alast);
        D.next_rebuild = (state == BOX_INIT ? 0 : k) + nK;
        D.builds += 1;
        D.built_at = state == BOX_INIT ? 0 : k;
        D.vmax2 = 0ull;
        D.amax2 = 0ull;
    } else if (review) {
        // the next step would build: postpone while the measured displacement
        // (gathered by the step pass this step) leaves room, by the same bound
        // as reuse_steps on what is left
        D.dmax2 = 0ull;
        const int64_t c = extend_steps(A.boxes[b], D, dmax2);
        if (c > 0) {
            D.next_rebuild = k + 1 + c;
            D.extends += 1;
        }
    }
    if (state == BOX_INIT) {
        D.state = K > 0 ? BOX_RUN : BOX_DONE;
        D.k = 0;
    } else if (state == BOX_RUN) {
        if (k >= K) D.state = BOX_DONE;
        else D.k = k + 1;
    }
    if (A.dyn_host) A.dyn_host[b] = D;
}

// ===========================================================================
// transfers between caller order (local ids) and slots
// ===========================================================================

// field codes: 0 r, 1 v, 2 m, 3 u (upload); download adds 4 rho, 5 rho_lr,
// 6 T, 7 p, 8 pco, 9 acc, 10 udot, 11 gradv, 12 jq, 13 P
__global__ __launch_bounds__(NT) void k_upload(ArenaView A, int off, int n, int field,
                                                const double *__restrict__ src)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const int s = off + i;
    const int k = A.id[s];
    switch (field) {
    case 0: A.r[s] = make_double2(src[2 * k], src[2 * k + 1]); break;
    case 1: {
        const double2 v = make_double2(src[2 * k], src[2 * k + 1]);
        A.v[s] = v;
        A.vp[s] = v;
    } break;
    case 2: A.m[s] = src[k]; break;
    case 3: A.u[s] = src[k]; A.up[s] = src[k]; break;
    default: break;
    }
}

__global__ __launch_bounds__(NT) void k_init_slots(ArenaView A, int off, int n)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const int s = off + i;
    A.id[s] = i;
    A.acc[s] = make_double2(0.0, 0.0);
    A.udot[s] = 0.0;
}

__global__ __launch_bounds__(NT) void k_download(ArenaView A, int box, int field,
                                                  double *__restrict__ dst)
{
    const BoxDesc &B = A.boxes[box];
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= B.n) return;
    const int s = B.part_off + i;
    const int k = A.id[s];
    switch (field) {
    case 0: { const double2 x = A.r[s]; dst[2 * k] = x.x; dst[2 * k + 1] = x.y; } break;
    case 1: { const double2 x = A.v[s]; dst[2 * k] = x.x; dst[2 * k + 1] = x.y; } break;
    case 2: dst[k] = A.m[s]; break;
    case 3: dst[k] = A.u[s]; break;
    case 4: dst[k] = A.rho[s].x; break;
    case 5: dst[k] = A.rho[s].y; break;
    case 6: dst[k] = A.T[s]; break;
    case 7: dst[k] = vdw_p(A.rho[s].x, A.T[s], B); break;
    case 8: { const double rl = A.rho[s].y; dst[k] = -B.a * rl * rl; } break;
    case 9: { const double2 x = A.acc[s]; dst[2 * k] = x.x; dst[2 * k + 1] = x.y; } break;
    case 10: dst[k] = A.udot[s]; break;
    case 11: { const double4 g = A.gradv[s];
               dst[4 * k] = g.x; dst[4 * k + 1] = g.y; dst[4 * k + 2] = g.z; dst[4 * k + 3] = g.w; } break;
    case 12: { const double2 q = A.jq[s]; dst[2 * k] = q.x; dst[2 * k + 1] = q.y; } break;
    case 13: {  // P = p_rev + p_rev_lr + pi_irr (spam_complete_force.py:178)
        const double rho = A.rho[s].x, rl = A.rho[s].y, r2 = rho * rho;
        double4 S;
        if (B.precision == SPH_FP32C) {  // the FP32C form keeps its stress record only
            const float4 f = A.S32[s];
            S = make_double4(f.x, f.y, f.z, f.w);
        } else {
            S = A.S[s];
        }
        const double pco = (B.terms & 2u) ? -B.a * rl * rl : 0.0;
        dst[4 * k] = S.x * r2 + pco;
        dst[4 * k + 1] = S.y * r2;
        dst[4 * k + 2] = S.y * r2;
        dst[4 * k + 3] = S.z * r2 + pco;
    } break;
    default: break;
    }
}

// ===========================================================================
// (a) pair export, separations, kernel and minimum image on demand
// ===========================================================================

// inv[local id] = slot
__global__ __launch_bounds__(NT) void k_inverse(const int *__restrict__ ids, int off, int n,
                                                 int *__restrict__ inv)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    inv[ids[off + i]] = off + i;
}

// Pairs (i, j), i < j, rsq < R2list: count per i (pass 0) or write the j's
// of i into its segment, sorted ascending (pass 1).  i runs over local ids;
// the cells are those of the build that precedes this call.
__global__ __launch_bounds__(NT) void k_pairs(ArenaView A, int box, const int *__restrict__ inv,
                                               unsigned *__restrict__ count,
                                               const unsigned *__restrict__ offs,
                                               int *__restrict__ jbuf, int pass)
{
    const BoxDesc &B = A.boxes[box];
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= B.n) return;
    const int s = inv[i];
    const double2 ri = A.r[s];
    unsigned c = 0;
    const unsigned o = pass ? offs[i] : 0u;
    for_each_cell(B, A.cell_start, A.pcell[s], [&](int t0, int t1, int, int) {
        for (int t = t0; t < t1; ++t) {
            const int j = A.id[t];
            if (j <= i) continue;
            const double2 rj = A.r[t];
            double dx = rj.x - ri.x, dy = rj.y - ri.y;
            min_image(dx, dy, B);
            if (rsq_exact(dx, dy) < B.R2list) {
                if (pass) jbuf[o + c] = j;
                ++c;
            }
        }
    });
    if (!pass) {
        count[i] = c;
        return;
    }
    int *seg = jbuf + o;  // insertion sort of this particle's partners
    for (unsigned a = 1; a < c; ++a) {
        const int x = seg[a];
        int q = (int)a - 1;
        while (q >= 0 && seg[q] > x) {
            seg[q + 1] = seg[q];
            --q;
        }
        seg[q + 1] = x;
    }
}

// Expand (offsets, j-lists) into int64 iap[k][2] (first cap pairs).
__global__ __launch_bounds__(NT) void k_pairs_emit(int n, const unsigned *__restrict__ offs,
                                                    const unsigned *__restrict__ count,
                                                    const int *__restrict__ jbuf,
                                                    long long *__restrict__ iap, long long cap)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const unsigned o = offs[i], c = count[i];
    for (unsigned q = 0; q < c; ++q) {
        const long long k = (long long)o + q;
        if (k < cap) {
            iap[2 * k] = i;
            iap[2 * k + 1] = jbuf[o + q];
        }
    }
}

// neighbour_list.py:63-83 on a given pair list (local ids, current state).
__global__ __launch_bounds__(NT) void k_separations(ArenaView A, int box, const int *__restrict__ inv,
                                                     long long np, const long long *__restrict__ iap,
                                                     double *__restrict__ drij, double *__restrict__ rij,
                                                     double *__restrict__ rsq, double *__restrict__ dv)
{
    const long long k = (long long)blockIdx.x * NT + threadIdx.x;
    if (k >= np) return;
    const BoxDesc &B = A.boxes[box];
    const int si = inv[iap[2 * k]], sj = inv[iap[2 * k + 1]];
    const double2 ri = A.r[si], rj = A.r[sj];
    double dx = rj.x - ri.x, dy = rj.y - ri.y;
    min_image(dx, dy, B);
    const double q = rsq_exact(dx, dy);
    if (drij) { drij[2 * k] = dx; drij[2 * k + 1] = dy; }
    if (rsq) rsq[k] = q;
    if (rij) rij[k] = sqrt(q);
    if (dv) {
        const double2 vi = A.v[si], vj = A.v[sj];
        dv[2 * k] = vj.x - vi.x;
        dv[2 * k + 1] = vj.y - vi.y;
    }
}

// spkernel.lucy_kernel (spkernel.py:86-118) for np pairs, dim 1..3; the
// gradient uses the factored form dW/dr / r = -12 q (1 - r/h)^2 / h^2.
__global__ __launch_bounds__(NT) void k_lucy(long long np, const double *__restrict__ r,
                                              const double *__restrict__ dx, double h, int dim,
                                              double *__restrict__ w, double *__restrict__ dwdx)
{
    const long long k = (long long)blockIdx.x * NT + threadIdx.x;
    if (k >= np) return;
    const double PI = 3.14159265358979323846;
    const double q = dim == 1 ? 5.0 / (4.0 * h) : (dim == 2 ? 5.0 / (PI * h * h) : 105.0 / (PI * 16.0 * (h * h * h)));
    const double rr = fabs(r[k]);
    double W = 0.0, F = 0.0;
    if (rr < h) {
        const double s = rr / h, t = 1.0 - s;
        W = q * (1.0 + 3.0 * s) * (t * t * t);
        // dwdx is 0 at r == 0 whatever dx says (spkernel.py:108-109)
        F = rr == 0.0 ? 0.0 : -12.0 * q * (t * t) / (h * h);
    }
    w[k] = W;
    for (int c = 0; c < dim; ++c) dwdx[dim * k + c] = F * dx[dim * k + c];
}

__global__ __launch_bounds__(NT) void k_min_image(long long np, double *__restrict__ dr, double Lx, double Ly)
{
    const long long k = (long long)blockIdx.x * NT + threadIdx.x;
    if (k >= np) return;
    double x = dr[2 * k], y = dr[2 * k + 1];
    x = x > Lx / 2. ? x - Lx : x;
    y = y > Ly / 2. ? y - Ly : y;
    x = x < -Lx / 2. ? x + Lx : x;
    y = y < -Ly / 2. ? y + Ly : y;
    dr[2 * k] = x;
    dr[2 * k + 1] = y;
}

// u = kB T - a rho for one box (T in caller order, or T0 when T is null).
__global__ __launch_bounds__(NT) void k_set_u(ArenaView A, int box, const double *__restrict__ T,
                                               double T0)
{
    const BoxDesc &B = A.boxes[box];
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= B.n) return;
    const int s = B.part_off + i;
    const double t = T ? T[A.id[s]] : T0;
    const double u = B.kB * t - B.a * A.rho[s].x;
    A.u[s] = u;
    A.up[s] = u;
}

// out3 = (sum m v^2/2, sum m u, total) over one box's state.
__global__ __launch_bounds__(NT) void k_energy_now(ArenaView A, int box, double *__restrict__ out3)
{
    const BoxDesc &B = A.boxes[box];
    double ke = 0.0, ue = 0.0;
    for (int i = threadIdx.x; i < B.n; i += NT) {
        const int s = B.part_off + i;
        const double2 v = A.v[s];
        const double m = A.m[s];
        ke += m * (0.5 * (v.x * v.x + v.y * v.y));
        ue += m * A.u[s];
    }
    __shared__ double rk[NT], ru[NT];
    rk[threadIdx.x] = ke;
    ru[threadIdx.x] = ue;
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            rk[threadIdx.x] += rk[threadIdx.x + off];
            ru[threadIdx.x] += ru[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out3[0] = rk[0];
        out3[1] = ru[0];
        out3[2] = rk[0] + ru[0];
    }
}

// ===========================================================================
// SPH scatter to a grid (interpolate.splash_map_2d) as a per-pixel gather
// ===========================================================================

// Particle record (x, y wrapped into the periods, w = a m / rho, h) and its
// cell key.
__global__ __launch_bounds__(NT) void k_splash_prep(long long n, const double *__restrict__ r,
                                                     const double *__restrict__ m,
                                                     const double *__restrict__ rho,
                                                     const double *__restrict__ a,
                                                     const double *__restrict__ h, SplashGrid G,
                                                     double4 *__restrict__ rec, unsigned *__restrict__ keys)
{
    const long long i = (long long)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    double x = r[2 * i], y = r[2 * i + 1];
    x -= G.Lx * floor(x / G.Lx);
    y -= G.Ly * floor(y / G.Ly);
    int cx = (int)(x / G.cwx), cy = (int)(y / G.cwy);
    cx = cx < 0 ? 0 : (cx >= G.ncx ? G.ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= G.ncy ? G.ncy - 1 : cy);
    rec[i] = make_double4(r[2 * i], r[2 * i + 1], a[i] * m[i] / rho[i], h[i]);
    keys[i] = (unsigned)(cy * G.ncx + cx);
}

__global__ __launch_bounds__(NT) void k_splash_gather(long long n, const unsigned *__restrict__ src,
                                                       const double4 *__restrict__ rec,
                                                       double4 *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    out[i] = rec[src[i]];
}

// cs[c] = first sorted particle with key >= c, c in [0, ncell]
__global__ __launch_bounds__(NT) void k_splash_cells(long long n, const unsigned *__restrict__ keys, int ncell,
                                                      int *__restrict__ cs)
{
    const long long i = (long long)blockIdx.x * NT + threadIdx.x;
    if (i > n) return;
    const int k = i == n ? ncell : (int)keys[i];
    const int kp = i == 0 ? -1 : (int)keys[i - 1];
    for (int c = kp + 1; c <= k; ++c) cs[c] = (int)i;
}

// One thread per pixel (i, j) at (x[i], y[j]); Z[i * ny + j] sums
// w_p W(s, h_p) over particles within the cutoff (minimum image with periods
// Lx, Ly, dr = (x_p - x, y - y_p) as interpolate.py:76).  Cells are at least
// the cutoff wide, so the 3x3 stencil of the pixel's cell holds them all
// (every cell when an axis has fewer than three).
__global__ __launch_bounds__(NT) void k_splash_pixels(int nx, int ny, const double *__restrict__ x,
                                                       const double *__restrict__ y, SplashGrid G,
                                                       const double4 *__restrict__ rec,
                                                       const int *__restrict__ cs, double *__restrict__ Z)
{
    const long long pix = (long long)blockIdx.x * NT + threadIdx.x;
    if (pix >= (long long)nx * ny) return;
    const int i = (int)(pix / ny), j = (int)(pix % ny);
    const double px = x[i], py = y[j];
    double wx = px - G.Lx * floor(px / G.Lx), wy = py - G.Ly * floor(py / G.Ly);
    int cx = (int)(wx / G.cwx), cy = (int)(wy / G.cwy);
    cx = cx < 0 ? 0 : (cx >= G.ncx ? G.ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= G.ncy ? G.ncy - 1 : cy);
    const int sx = G.ncx >= 3 ? 1 : 0, sy = G.ncy >= 3 ? 1 : 0;
    const int x0 = sx ? -1 : 0, x1 = sx ? 1 : G.ncx - 1, y0 = sy ? -1 : 0, y1 = sy ? 1 : G.ncy - 1;
    const double c2 = G.cutoff * G.cutoff;
    const double hLx = G.Lx / 2., hLy = G.Ly / 2.;
    double z = 0.0;
    for (int oy = y0; oy <= y1; ++oy) {
        const int yy = sy ? wrapi(cy + oy, G.ncy) : oy;
        for (int ox = x0; ox <= x1; ++ox) {
            const int xx = sx ? wrapi(cx + ox, G.ncx) : ox;
            const int c = yy * G.ncx + xx;
            for (int t = cs[c]; t < cs[c + 1]; ++t) {
                const double4 p = rec[t];
                double dx = p.x - px, dy = py - p.y;
                dx = dx > hLx ? dx - G.Lx : dx;  // neighbour_list.py:110-123 order
                dy = dy > hLy ? dy - G.Ly : dy;
                dx = dx < -hLx ? dx + G.Lx : dx;
                dy = dy < -hLy ? dy + G.Ly : dy;
                const double rsq = rsq_exact(dx, dy);
                if (rsq < c2) {
                    const double sr = sqrt(rsq);
                    if (sr < p.w) {  // spkernel.lucy_w, 2-D
                        const double q = sr / p.w, tt = 1.0 - q;
                        z += p.z * ((5.0 / (SPH_PI * p.w * p.w)) * (1.0 + 3.0 * q) * (tt * tt * tt));
                    }
                }
            }
        }
    }
    Z[pix] = z;
}

// ===========================================================================
// host launchers
// ===========================================================================

static inline dim3 grid_for(long long n) { return dim3((unsigned)((n + NT - 1) / NT > 0 ? (n + NT - 1) / NT : 1)); }
static inline dim3 grid_for_bs(long long n, int bs)
{
    return dim3((unsigned)((n + bs - 1) / bs > 0 ? (n + bs - 1) / bs : 1));
}
static inline int pass_block_of(const ArenaView &A)
{
    return (A.pass_block == 128 || A.pass_block == 512) ? A.pass_block : NT;
}

void launch_slot_box(const ArenaView &A, hipStream_t st)
{
    if (A.n > 0) hipLaunchKernelGGL(k_slot_box, grid_for(A.n), dim3(NT), 0, st, A);
}
void launch_build_plan(const ArenaView &A, int force_mask, hipStream_t st)
{
    hipLaunchKernelGGL(k_build_plan, dim3(1), dim3(PLAN_THREADS), 0, st, A, force_mask);
}
// grid of the kernels over building tiles: at most one wave per tile of the
// slots in use, a multiple of 8 (an equal share per XCD), 8 to 2048 blocks
static inline dim3 tile_grid(const ArenaView &A)
{
    long long nb = ((long long)A.n + NT - 1) / NT;
    nb = (nb + 7) / 8 * 8;
    nb = nb < 8 ? 8 : (nb > 2048 ? 2048 : nb);
    return dim3((unsigned)nb);
}
void launch_keys(const ArenaView &A, int force_mask, hipStream_t st)
{
    hipLaunchKernelGGL(k_keys, tile_grid(A), dim3(NT), 0, st, A, force_mask);
}
void launch_permute(const ArenaView &A, int force_mask, int cells, hipStream_t st)
{
    hipLaunchKernelGGL(k_permute, tile_grid(A), dim3(NT), 0, st, A, force_mask, cells);
}
void launch_box_sort(const ArenaView &A, int force_mask, int nbox, hipStream_t st)
{
    if (nbox > 0) hipLaunchKernelGGL(k_box_sort, dim3(nbox), dim3(BS_THREADS), 0, st, A, force_mask, nbox);
}
// one block sorts a box: beyond ~16k slots its sequential walk outlasts the
// radix sort (measured on the served workload, boxes up to 40k slots)
bool box_sort_fits(int ncells, int n) { return ncells <= BS_MAX_CELLS && n <= 16384; }
void launch_build_lists(const ArenaView &A, int force_mask, int fp32, hipStream_t st)
{
    hipLaunchKernelGGL(k_build_lists, tile_grid(A), dim3(NT), 0, st, A, force_mask, fp32);
}
// unroll: entries in flight per lane (1, 2 or 4; anything else = default)
#define SPH_DISPATCH(K64, K32, ...)                                                              \
    do {                                                                                         \
        const int bs_ = pass_block_of(A);                                                        \
        const dim3 g = grid_for_bs(A.n, bs_), b(bs_);                                            \
        if (fp32) {                                                                              \
            if (unroll == 1) hipLaunchKernelGGL(K32<1>, g, b, 0, st, __VA_ARGS__);               \
            else if (unroll == 2) hipLaunchKernelGGL(K32<2>, g, b, 0, st, __VA_ARGS__);          \
            else hipLaunchKernelGGL(K32<4>, g, b, 0, st, __VA_ARGS__);                           \
        } else {                                                                                 \
            if (unroll == 1) hipLaunchKernelGGL(K64<1>, g, b, 0, st, __VA_ARGS__);               \
            else if (unroll == 4) hipLaunchKernelGGL(K64<4>, g, b, 0, st, __VA_ARGS__);          \
            else hipLaunchKernelGGL(K64<2>, g, b, 0, st, __VA_ARGS__);                           \
        }                                                                                        \
    } while (0)

void launch_density(const ArenaView &A, int fp32, int state_mask, int unroll, hipStream_t st)
{
    SPH_DISPATCH(k_density64, k_density32, A, state_mask);
}
void launch_grad(const ArenaView &A, int fp32, int state_mask, int diag, int unroll, hipStream_t st)
{
    SPH_DISPATCH(k_grad64, k_grad32, A, state_mask, diag);
}
void launch_rates(const ArenaView &A, int fp32, int state_mask, int unroll, int step, hipStream_t st)
{
    SPH_DISPATCH(k_rates64, k_rates32, A, state_mask, step);
}
void launch_leapfrog(const ArenaView &A, int fp32, hipStream_t st)
{
    hipLaunchKernelGGL(k_leapfrog, grid_for(A.n), dim3(NT), 0, st, A, fp32);
}
void launch_box_step(const ArenaView &A, int nbox, hipStream_t st)
{
    if (nbox > 0) hipLaunchKernelGGL(k_box_step, dim3(nbox), dim3(NT_BOX), 0, st, A, nbox);
}
void launch_upload(const ArenaView &A, int off, int n, int field, const double *src, hipStream_t st)
{
    hipLaunchKernelGGL(k_upload, grid_for(n), dim3(NT), 0, st, A, off, n, field, src);
}
void launch_init_slots(const ArenaView &A, int off, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_slots, grid_for(n), dim3(NT), 0, st, A, off, n);
}
void launch_download(const ArenaView &A, int box, int n, int field, double *dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_download, grid_for(n), dim3(NT), 0, st, A, box, field, dst);
}
void launch_inverse(const int *ids, int off, int n, int *inv, hipStream_t st)
{
    hipLaunchKernelGGL(k_inverse, grid_for(n), dim3(NT), 0, st, ids, off, n, inv);
}
void launch_pairs(const ArenaView &A, int box, int n, const int *inv, unsigned *count,
                  const unsigned *offs, int *jbuf, int pass, hipStream_t st)
{
    hipLaunchKernelGGL(k_pairs, grid_for(n), dim3(NT), 0, st, A, box, inv, count, offs, jbuf, pass);
}
void launch_pairs_emit(int n, const unsigned *offs, const unsigned *count, const int *jbuf,
                       long long *iap, long long cap, hipStream_t st)
{
    hipLaunchKernelGGL(k_pairs_emit, grid_for(n), dim3(NT), 0, st, n, offs, count, jbuf, iap, cap);
}
void launch_separations(const ArenaView &A, int box, const int *inv, long long np,
                        const long long *iap, double *drij, double *rij, double *rsq, double *dv,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_separations, grid_for(np), dim3(NT), 0, st, A, box, inv, np, iap, drij, rij, rsq, dv);
}
void launch_splash(long long n, const double *r, const double *m, const double *rho, const double *a,
                   const double *h, const SplashGrid &G, double4 *rec, double4 *rec_sorted, unsigned *keys,
                   unsigned *k_out, unsigned *v_out, unsigned *k_alt, unsigned *v_alt, void *sort_tmp, int key_bits,
                   int *cs, int nx, int ny, const double *x, const double *y, double *Z, hipStream_t st)
{
    if (n > 0) {
        hipLaunchKernelGGL(k_splash_prep, grid_for(n), dim3(NT), 0, st, n, r, m, rho, a, h, G, rec, keys);
        radix_sort_pairs(keys, k_out, v_out, k_alt, v_alt, (int)n, key_bits, sort_tmp, st);
        hipLaunchKernelGGL(k_splash_gather, grid_for(n), dim3(NT), 0, st, n, v_out, rec, rec_sorted);
    }
    hipLaunchKernelGGL(k_splash_cells, grid_for(n + 1), dim3(NT), 0, st, n, k_out, G.ncx * G.ncy, cs);
    const long long np = (long long)nx * ny;
    hipLaunchKernelGGL(k_splash_pixels, grid_for(np), dim3(NT), 0, st, nx, ny, x, y, G, rec_sorted, cs, Z);
}
void launch_lucy(long long np, const double *r, const double *dx, double h, int dim, double *w,
                 double *dwdx, hipStream_t st)
{
    hipLaunchKernelGGL(k_lucy, grid_for(np), dim3(NT), 0, st, np, r, dx, h, dim, w, dwdx);
}
void launch_min_image(long long np, double *dr, double Lx, double Ly, hipStream_t st)
{
    hipLaunchKernelGGL(k_min_image, grid_for(np), dim3(NT), 0, st, np, dr, Lx, Ly);
}
void launch_set_u(const ArenaView &A, int box, int n, const double *T, double T0, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_u, grid_for(n), dim3(NT), 0, st, A, box, T, T0);
}
void launch_energy_now(const ArenaView &A, int box, double *out3, hipStream_t st)
{
    hipLaunchKernelGGL(k_energy_now, dim3(1), dim3(NT), 0, st, A, box, out3);
}
void launch_fp32_accumulate(long long n, const double *x, int unroll, double *out2, hipStream_t st)
{
    hipLaunchKernelGGL(k_fp32_accumulate, dim3(1), dim3(64), 0, st, n, x, unroll, out2);
}

}  // namespace sph
