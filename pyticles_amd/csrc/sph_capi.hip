// sph_capi.hip -- host runtime and C ABI (include/pyticles_hip.h).
//
// One sph_ctx = one device + one HIP stream.  All particle state lives in an
// Arena (device-resident structure of arrays, see sph_common.hpp); a sph_box
// is an arena holding a single box.  Every pass is launched on the context
// stream; host synchronisation happens only where a result must come back.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pyticles_hip.h"
#include "sph_common.hpp"
#include "sph_kernels.hpp"
#include "sph_sort.hpp"

using namespace sph;

namespace {

thread_local std::string g_err = "no error";

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(SPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(SPH_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                       " bytes failed: " + hipGetErrorString(e));
    }
    return SPH_OK;
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

int bits_for(int nkeys)
{
    int b = 1;
    while ((1ll << b) < (long long)nkeys) ++b;
    return b;
}

}  // namespace

// ---- fixed scratch of the allocation-free rendering entry point -------------
// sph_splash_map_2d_noalloc keeps its particle records, sort keys, cell starts
// and radix-sort counters here instead of allocating: a file-scope device
// array, one copy per GPU, held until the process ends.  Host code never
// uses the name itself (on the host it is a stand-in, not a device address):
// g_noalloc_scratch[device] holds the address hipGetSymbolAddress gives for
// it on that device, taken once per device when a context is created.
namespace {
constexpr int64_t NOALLOC_MAX_PARTICLES = 65536;
constexpr int NOALLOC_MAX_CELLS = 65536;
constexpr size_t NOALLOC_SORT_BYTES = radix_sort_temp_bytes(NOALLOC_MAX_PARTICLES);  // 129 KB
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NOALLOC_REC_BYTES = align256(sizeof(double4) * NOALLOC_MAX_PARTICLES);
constexpr size_t NOALLOC_KEY_BYTES = align256(sizeof(unsigned) * 5 * NOALLOC_MAX_PARTICLES);
constexpr size_t NOALLOC_CS_BYTES = align256(sizeof(int) * (NOALLOC_MAX_CELLS + 1));
constexpr size_t NOALLOC_SCRATCH_BYTES = 2 * NOALLOC_REC_BYTES + NOALLOC_KEY_BYTES + NOALLOC_CS_BYTES +
                                         NOALLOC_SORT_BYTES;
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_scratch[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// The device address of `scratch` on the current device (once per device).
static int noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(SPH_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_scratch[device]) return SPH_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SPH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_scratch[device] = p;
    return SPH_OK;
}

struct sph_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int timing = 0;
    // scratch for the stateless helpers (kernel, minimum image)
    double *d_scratch = nullptr;
    size_t scratch_bytes = 0;

    int scratch(size_t bytes)
    {
        if (bytes <= scratch_bytes) return SPH_OK;
        dfree(d_scratch);
        scratch_bytes = 0;
        int rc = dalloc((char **)&d_scratch, bytes);
        if (rc) return rc;
        scratch_bytes = bytes;
        return SPH_OK;
    }
};

namespace {

enum Timer { T_SORT = 0, T_DENS, T_GRAD, T_RATES, T_PAIRS, T_INT, T_N };

// Neighbour-list entries per slot in an arena: room for the densest state a
// served box reaches (liquid at rho ~ 1.8 within r_cut + skin = 4.5 holds
// ~115 neighbours).
static constexpr int ARENA_LIST_ENTRIES = 192;

// Device bytes Arena::alloc takes for cap slots (kept in step with alloc).
static size_t arena_bytes(size_t cap, int prec, int lcap)
{
    const size_t d2 = sizeof(double2), d1 = sizeof(double), i1 = sizeof(int);
    size_t b = 0;
    b += cap * (4 * d2 + 4 * d1 + i1);          // r v vp acc, u m up udot, id
    b += cap * 2 * d2;                          // second buffers of r, vp
    b += cap * (4 * d2 + 4 * d1 + i1);          // sorted scratch of the same
    b += cap * (2 * i1 + d2 + sizeof(int2));    // pcell slot_box, r_ref, cnt
    b += (cap + 63) / 64 * 64 * (size_t)lcap * sizeof(unsigned);  // nbr, in tiles of 64 slots
    b += cap * (3 * d2 + 2 * d1 + 2 * sizeof(double4));  // rho Q jq, T e, S gradv
    b += cap * (2 * d2 + d1);                             // vt, aco, uco
    if (prec == SPH_FP32C)  // rf S32, rlo, g32 q32, second buffers of rf, rlo
        b += cap * (2 * sizeof(float4) + sizeof(unsigned) + sizeof(GRec32) + sizeof(QRec32) + sizeof(float4) +
                    sizeof(unsigned));
    b += cap * 5 * sizeof(unsigned);             // sort keys / values
    b += ((cap + 63) / 64 + 1) * i1;             // building tiles
    b += (cap + 63) / 64 * sizeof(double);       // per-wave displacement
    b += (cap + 63) / 64 * sizeof(double2);      // per-wave build speed, acceleration
    b += (cap + 63) / 64 * sizeof(double);       // per-wave energy
    b += radix_sort_temp_bytes((int)cap);
    return b;
}

// Step count of an open run (sph_box_advance): no step reaches it, so the
// run keeps kicking by dt until sph_box_sync sets its closing step.
static constexpr int64_t OPEN_STEPS = (int64_t)1 << 60;

struct Arena {
    sph_ctx *ctx = nullptr;
    int precision = SPH_FP64;
    int cap = 0;   // slot capacity
    int hwm = 0;   // slots in use (end of the last box)

    int lcap = 0;  // neighbour-list entries per slot

    // device state (cap slots)
    double2 *r = nullptr, *v = nullptr, *vp = nullptr, *acc = nullptr;
    // second buffers: the step pass writes the stepped r, vp (rf, rlo) there
    // and swap_state() makes them current after every launched step
    double2 *r_nx = nullptr, *vp_nx = nullptr;
    float4 *rf_nx = nullptr;
    unsigned *rlo_nx = nullptr;
    int parity = 0;  // which set of buffers is current (swap_state)
    double *u = nullptr, *m = nullptr, *up = nullptr, *udot = nullptr;
    int *id = nullptr;
    double2 *r_s = nullptr, *v_s = nullptr, *vp_s = nullptr, *acc_s = nullptr;
    double *u_s = nullptr, *m_s = nullptr, *up_s = nullptr, *udot_s = nullptr;
    int *id_s = nullptr, *pcell = nullptr, *slot_box = nullptr;
    double2 *r_ref = nullptr;
    unsigned *nbr = nullptr;
    int2 *cnt = nullptr;
    double2 *rho = nullptr, *Q = nullptr, *jq = nullptr;
    double *T = nullptr, *e = nullptr;
    double4 *S = nullptr, *gradv = nullptr;
    float4 *rf = nullptr, *S32 = nullptr;
    unsigned *rlo = nullptr;
    GRec32 *g32 = nullptr;
    QRec32 *q32 = nullptr;
    double2 *vt = nullptr, *aco = nullptr;
    double *uco = nullptr;
    unsigned *keys_in = nullptr, *k_out = nullptr, *v_out = nullptr, *k_alt = nullptr, *v_alt = nullptr;
    int *plan_n = nullptr, *box_coff = nullptr;  // box_coff sized box_cap
    int *tile_list = nullptr;                     // building tiles, one per 64 slots
    double *wdisp = nullptr;                      // one per 64 slots
    double2 *wbuild = nullptr;                    // one per 64 slots
    double *wenergy = nullptr;                    // one per 64 slots
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int *cell_start = nullptr;
    int cell_cap = 0;

    // box / segment tables
    std::vector<BoxDesc> boxes;
    std::vector<int> box_live;
    std::vector<SegDesc> segs;
    BoxDesc *d_boxes = nullptr;
    BoxDyn *d_dyn = nullptr;
    SegDesc *d_segs = nullptr;
    int box_cap = 0, seg_cap = 0;
    int nkeys = 1;
    BoxDyn *h_dyn = nullptr;  // pinned mirror for polling
    // single boxes: the states k_box_step leaves in host-coherent memory at
    // the end of every step, and their device address; host_dyn_ok while
    // they equal d_dyn (the last launch that can change a box state was a
    // step: launch_steps sets it, every other writer of d_dyn clears it)
    bool keep_host_dyn = false;
    BoxDyn *h_stepped = nullptr, *d_stepped = nullptr;
    bool host_dyn_ok = false;

    // staging
    double *d_stage = nullptr;
    size_t stage_bytes = 0;
    double *h_stage = nullptr;  // pinned
    size_t h_stage_bytes = 0;

    // asynchronous snapshots: state copied on the device in stream order,
    // then to the host on a copy stream
    double *d_snap = nullptr;
    size_t snap_cap = 0;
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_snap_ready = nullptr, ev_snap_done = nullptr;
    bool snap_pending = false;
    double *d_e3 = nullptr;

    // pair export buffers
    int *d_inv = nullptr;
    unsigned *d_cnt = nullptr, *d_off = nullptr, *d_grand = nullptr;
    int *d_jbuf = nullptr;
    long long *d_iap = nullptr;
    size_t jbuf_cap = 0, iap_cap = 0;
    void *scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;

    // timing
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<size_t, size_t>>> ev_marks;
    double t_acc[T_N] = {0, 0, 0, 0, 0, 0};
    double steps_done = 0;
    bool stepped = false;  // a box (single-box use) has taken a sph_step
    // open run of a single box (sph_box_advance / sph_box_sync): the box
    // runs with an unbounded step count, so no step is its closing one
    bool open_run = false;
    int64_t open_plain = 0;  // next steps known not to build
    // the single box's lists are still valid for its state: its last run
    // ended without a health bit and no upload, parameter change, relayout
    // or out-of-run list build came after it (run_start)
    bool sched_kept = false;
    // list entry width for new boxes and set_list_bits (16: 16-bit offsets,
    // a box falling back to 32-bit slots when its offsets do not fit)
    int list_bits = 16;
    int64_t released = 0;  // boxes switched to 32-bit entries after a held build (read_dyn)

    // captured graphs of one arena step without / with a list build
    // step graphs: [0] one step without the build stage, [1] one step with
    // it, [2 + j] 2^(j+1) steps without it (a box's run between two builds
    // replays the largest that fit: each graph launch costs ~1-2 us of idle
    // GPU between graphs)
    static constexpr int PLAIN_LOG2 = 3;
    hipGraphExec_t graph[2][2 + PLAIN_LOG2] = {};  // per parity (the buffers a step starts from)
    // What the captured launches bake in: the view (pointers, slot, key and
    // segment counts), the box count, the sort path and the pass depths.  A
    // relayout that leaves it unchanged (a box's terms or constants changed:
    // the kernels read those from the box table) keeps the graphs.
    struct GraphSig {
        ArenaView view;
        int nbox, small, unroll, unroll_rates, fuse_step;
    };
    GraphSig graph_sig;
    bool graph_sig_set = false;
    int use_graph = 1;
    int always_build = 0;  // arenas: build stage every step (per-box schedule on device)
    bool small_boxes = false;  // all live boxes fit the per-box sort (set by relayout)
    int unroll = 0;        // list entries in flight per lane in the pair passes (0 = default)
    int pass_block = 256;  // work-items per block of the pair passes (128, 256, 512)
    int unroll_rates = 0;  // the same for the rate pass alone (0 = unroll)
    int fuse_step = 1;     // the leapfrog step inside the rate pass (0: a launch of its own)
    int rates_unroll() const { return unroll_rates ? unroll_rates : unroll; }

    bool fp32() const { return precision == SPH_FP32C; }

    ArenaView view() const
    {
        ArenaView A;
        memset(&A, 0, sizeof(A));
        A.r = r; A.v = v; A.vp = vp; A.acc = acc; A.u = u; A.m = m; A.up = up; A.udot = udot; A.id = id;
        A.r_nx = r_nx; A.vp_nx = vp_nx; A.rf_nx = rf_nx; A.rlo_nx = rlo_nx;
        A.r_s = r_s; A.v_s = v_s; A.vp_s = vp_s; A.acc_s = acc_s;
        A.u_s = u_s; A.m_s = m_s; A.up_s = up_s; A.udot_s = udot_s; A.id_s = id_s;
        A.pcell = pcell; A.r_ref = r_ref; A.slot_box = slot_box; A.nbr = nbr; A.cnt = cnt;
        A.rho = rho; A.T = T; A.S = S; A.Q = Q; A.gradv = gradv; A.jq = jq; A.e = e;
        A.rf = rf; A.rlo = rlo; A.S32 = S32; A.g32 = g32; A.q32 = q32;
        A.vt = vt; A.aco = aco; A.uco = uco;
        A.cell_start = cell_start;
        A.keys = k_out; A.vals = v_out; A.keys_in = keys_in;
        A.plan_n = plan_n; A.box_coff = box_coff; A.tile_list = tile_list; A.wdisp = wdisp; A.wbuild = wbuild; A.wenergy = wenergy;
        A.boxes = d_boxes; A.dyn = d_dyn; A.dyn_host = d_stepped; A.segs = d_segs;
        A.nseg = (int)segs.size();
        A.n = hwm;
        A.nkeys = nkeys;
        A.cap = cap;
        A.lcap = lcap;
        A.pass_block = pass_block;
        return A;
    }

    int alloc(int capacity, int prec, int list_entries)
    {
        cap = capacity;
        precision = prec;
        lcap = list_entries;
        if (lcap <= 0 || lcap % 8) return fail(SPH_ERR_ARG, "list entries must be a positive multiple of 8");
        // the pair walks gather neighbour records (up to 32 bytes) through
        // 32-bit byte offsets
        if (cap > (1 << 27)) return fail(SPH_ERR_ARG, "arena capacity above 2^27 slots");
        int rc = 0;
        rc |= dalloc(&r, cap); rc |= dalloc(&v, cap); rc |= dalloc(&vp, cap); rc |= dalloc(&acc, cap);
        rc |= dalloc(&r_nx, cap); rc |= dalloc(&vp_nx, cap);
        rc |= dalloc(&u, cap); rc |= dalloc(&m, cap); rc |= dalloc(&up, cap); rc |= dalloc(&udot, cap);
        rc |= dalloc(&id, cap);
        rc |= dalloc(&r_s, cap); rc |= dalloc(&v_s, cap); rc |= dalloc(&vp_s, cap); rc |= dalloc(&acc_s, cap);
        rc |= dalloc(&u_s, cap); rc |= dalloc(&m_s, cap); rc |= dalloc(&up_s, cap); rc |= dalloc(&udot_s, cap);
        rc |= dalloc(&id_s, cap); rc |= dalloc(&pcell, cap); rc |= dalloc(&slot_box, cap);
        rc |= dalloc(&r_ref, cap); rc |= dalloc(&cnt, cap);
        rc |= dalloc(&nbr, (size_t)(cap + 63) / 64 * 64 * lcap);  // tiles of 64 slots
        rc |= dalloc(&rho, cap); rc |= dalloc(&Q, cap); rc |= dalloc(&jq, cap);
        rc |= dalloc(&T, cap); rc |= dalloc(&e, cap);
        rc |= dalloc(&S, cap); rc |= dalloc(&gradv, cap);
        rc |= dalloc(&vt, cap); rc |= dalloc(&aco, cap); rc |= dalloc(&uco, cap);
        if (prec == SPH_FP32C) {
            rc |= dalloc(&rf, cap); rc |= dalloc(&rlo, cap); rc |= dalloc(&S32, cap);
            rc |= dalloc(&rf_nx, cap); rc |= dalloc(&rlo_nx, cap);
            rc |= dalloc(&g32, cap); rc |= dalloc(&q32, cap);
        }
        rc |= dalloc(&keys_in, cap); rc |= dalloc(&k_out, cap); rc |= dalloc(&v_out, cap);
        rc |= dalloc(&k_alt, cap); rc |= dalloc(&v_alt, cap);
        sort_tmp_bytes = radix_sort_temp_bytes(cap);
        rc |= dalloc((char **)&sort_tmp, sort_tmp_bytes);
        rc |= dalloc(&d_e3, 3);
        rc |= dalloc(&plan_n, 4);
        rc |= dalloc(&tile_list, (size_t)(cap + 63) / 64 + 1);
        rc |= dalloc(&wdisp, (size_t)(cap + 63) / 64);
        rc |= dalloc(&wbuild, (size_t)(cap + 63) / 64);
        rc |= dalloc(&wenergy, (size_t)(cap + 63) / 64);
        if (rc) return SPH_ERR_NOMEM;
        // zero-fill the state so that unused slots hold finite values
        size_t z2 = sizeof(double2) * cap, z1 = sizeof(double) * cap;
        if (hipMemset(r, 0, z2) || hipMemset(v, 0, z2) || hipMemset(vp, 0, z2) || hipMemset(u, 0, z1) ||
            hipMemset(r_nx, 0, z2) || hipMemset(vp_nx, 0, z2) ||
            hipMemset(up, 0, z1) || hipMemset(m, 0, z1) || hipMemset(id, 0, sizeof(int) * cap) ||
            hipMemset(rho, 0, z2) || hipMemset(T, 0, z1) || hipMemset(S, 0, sizeof(double4) * cap) ||
            hipMemset(Q, 0, z2) || hipMemset(acc, 0, z2) || hipMemset(udot, 0, z1) ||
            hipMemset(gradv, 0, sizeof(double4) * cap) || hipMemset(jq, 0, z2) || hipMemset(e, 0, z1) ||
            hipMemset(r_ref, 0, z2) || hipMemset(cnt, 0, sizeof(int2) * cap) || hipMemset(vt, 0, z2) ||
            hipMemset(aco, 0, z2) || hipMemset(uco, 0, z1))
            return fail(SPH_ERR_HIP, "hipMemset failed");
        // the context stream does not synchronise with the null stream
        HIPCHK(hipDeviceSynchronize());
        segs.clear();
        segs.push_back(SegDesc{0, cap, -1, 0});
        return ensure_tables(4);
    }

    void release()
    {
        invalidate_graph();
        dfree(r); dfree(v); dfree(vp); dfree(acc); dfree(u); dfree(m); dfree(up); dfree(udot); dfree(id);
        dfree(r_nx); dfree(vp_nx); dfree(rf_nx); dfree(rlo_nx);
        dfree(r_s); dfree(v_s); dfree(vp_s); dfree(acc_s); dfree(u_s); dfree(m_s); dfree(up_s); dfree(udot_s);
        dfree(id_s); dfree(pcell); dfree(slot_box); dfree(r_ref); dfree(nbr); dfree(cnt);
        dfree(rho); dfree(Q); dfree(jq); dfree(T); dfree(e); dfree(S); dfree(gradv);
        dfree(rf); dfree(rlo); dfree(S32); dfree(g32); dfree(q32);
        dfree(vt); dfree(aco); dfree(uco);
        dfree(keys_in); dfree(k_out); dfree(v_out); dfree(k_alt); dfree(v_alt);
        dfree(plan_n); dfree(box_coff); dfree(tile_list); dfree(wdisp); dfree(wbuild); dfree(wenergy);
        if (sort_tmp) (void)hipFree(sort_tmp);
        sort_tmp = nullptr;
        if (snap_pending) (void)hipEventSynchronize(ev_snap_done);
        snap_pending = false;
        if (ev_snap_ready) (void)hipEventDestroy(ev_snap_ready);
        if (ev_snap_done) (void)hipEventDestroy(ev_snap_done);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        ev_snap_ready = ev_snap_done = nullptr;
        copy_stream = nullptr;
        dfree(d_snap);
        snap_cap = 0;
        dfree(cell_start); dfree(d_boxes); dfree(d_dyn); dfree(d_segs); dfree(d_stage); dfree(d_e3);
        if (h_stage) (void)hipHostFree(h_stage);
        h_stage = nullptr;
        h_stage_bytes = 0;
        dfree(d_inv); dfree(d_cnt); dfree(d_off); dfree(d_grand); dfree(d_jbuf); dfree(d_iap);
        if (scan_tmp) (void)hipFree(scan_tmp);
        scan_tmp = nullptr;
        if (h_dyn) (void)hipHostFree(h_dyn);
        h_dyn = nullptr;
        if (h_stepped) (void)hipHostFree(h_stepped);
        h_stepped = d_stepped = nullptr;
        for (auto &ev : ev_pool) (void)hipEventDestroy(ev);
        ev_pool.clear();
    }

    int ensure_tables(int nbox)
    {
        if (nbox > box_cap) {
            host_dyn_ok = false;
            int nc = std::max(nbox, box_cap * 2);
            BoxDesc *nb = nullptr;
            BoxDyn *nd = nullptr;
            int *nco = nullptr;
            if (dalloc(&nb, nc) || dalloc(&nd, nc) || dalloc(&nco, nc)) return SPH_ERR_NOMEM;
            HIPCHK(hipMemsetAsync(nd, 0, sizeof(BoxDyn) * nc, ctx->stream));
            if (box_cap > 0) {
                HIPCHK(hipMemcpyAsync(nd, d_dyn, sizeof(BoxDyn) * box_cap, hipMemcpyDeviceToDevice, ctx->stream));
            }
            HIPCHK(hipStreamSynchronize(ctx->stream));
            dfree(d_boxes);
            dfree(d_dyn);
            dfree(box_coff);
            d_boxes = nb;
            d_dyn = nd;
            box_coff = nco;
            BoxDyn *nh = nullptr;
            HIPCHK(hipHostMalloc((void **)&nh, sizeof(BoxDyn) * nc, hipHostMallocDefault));
            if (h_dyn) (void)hipHostFree(h_dyn);
            h_dyn = nh;
            if (keep_host_dyn) {
                BoxDyn *ns = nullptr, *nsd = nullptr;
                HIPCHK(hipHostMalloc((void **)&ns, sizeof(BoxDyn) * nc, hipHostMallocCoherent));
                memset(ns, 0, sizeof(BoxDyn) * nc);
                if (hipHostGetDevicePointer((void **)&nsd, ns, 0) != hipSuccess) {
                    (void)hipHostFree(ns);
                    return fail(SPH_ERR_HIP, "hipHostGetDevicePointer failed");
                }
                if (h_stepped) (void)hipHostFree(h_stepped);
                h_stepped = ns;
                d_stepped = nsd;
            }
            box_cap = nc;
            invalidate_graph();
        }
        return SPH_OK;
    }

    void invalidate_graph()
    {
        graph_sig_set = false;
        bool any = false;
        for (auto &gp : graph)
            for (auto g : gp) any = any || g;
        if (!any) return;
        (void)hipStreamSynchronize(ctx->stream);  // no replay of them is in flight
        for (auto &gp : graph)
            for (auto &g : gp) {
                if (g) (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }

    GraphSig signature() const
    {
        GraphSig sg;
        memset(&sg, 0, sizeof(sg));
        sg.view = view();
        if (parity) {  // the same graphs serve both parities: compare one order
            std::swap(sg.view.r, sg.view.r_nx);
            std::swap(sg.view.vp, sg.view.vp_nx);
            std::swap(sg.view.rf, sg.view.rf_nx);
            std::swap(sg.view.rlo, sg.view.rlo_nx);
        }
        sg.nbox = (int)boxes.size();
        sg.small = small_boxes ? 1 : 0;
        sg.unroll = unroll;
        sg.unroll_rates = unroll_rates;
        sg.fuse_step = fuse_step;
        return sg;
    }

    // Drop the captured graphs if what they bake in has changed since they
    // were captured.
    void check_graph_sig()
    {
        const GraphSig now = signature();
        if (graph_sig_set && memcmp(&now, &graph_sig, sizeof(now)) == 0) return;
        invalidate_graph();
        graph_sig = now;
        graph_sig_set = true;
    }

    // The stepped state becomes the current one: swap the buffer pairs the
    // step pass reads and writes (host pointers only; stream order puts every
    // later launch after the step).
    void swap_state()
    {
        std::swap(r, r_nx);
        std::swap(vp, vp_nx);
        std::swap(rf, rf_nx);
        std::swap(rlo, rlo_nx);
        parity ^= 1;
    }

    // Capture and instantiate step graph idx for the current parity: [0] one
    // step without the build stage, [1] one step with it, [1 + lg] 2^lg steps
    // without it (the buffers swapping between the captured steps).
    int capture_graph(int idx)
    {
        hipGraphExec_t &g = graph[parity][idx];
        if (g) return SPH_OK;
        check_graph_sig();
        const bool with_build = idx == 1;
        const int per = idx >= 2 ? 1 << (idx - 1) : 1;
        hipGraph_t gr;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        for (int j = 0; j < per; ++j) {
            enqueue_step(with_build);
            swap_state();
        }
        if (per & 1) swap_state();  // nothing ran: back to the buffers that hold the state
        HIPCHK(hipStreamEndCapture(ctx->stream, &gr));
        hipError_t e1 = hipGraphInstantiate(&g, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e1 != hipSuccess) {
            g = nullptr;
            return fail(SPH_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e1));
        }
        return SPH_OK;
    }

    // Set-up time capture of every step graph a single box replays, so no
    // capture or instantiation falls inside a stepping call.
    int prepare_graphs()
    {
        if (!use_graph || ctx->timing) return SPH_OK;  // timed runs launch directly
        HIPCHK(hipSetDevice(ctx->device));
        // both parities: a step may start from either set of buffers (the
        // swaps around the second set move pointers only)
        for (int pass = 0; pass < 2; ++pass) {
            int rc = SPH_OK;
            for (int idx = 0; idx < 2 + PLAIN_LOG2 && !rc; ++idx) rc = capture_graph(idx);
            swap_state();
            if (rc) {
                if (pass == 0) swap_state();
                return rc;
            }
        }
        return SPH_OK;
    }

    // Recompute key bases in slot order, the high-water mark, and upload the
    // box and segment tables.
    int relayout()
    {
        int key = 0;
        hwm = 0;
        sched_kept = false;  // slots, constants or boxes change
        for (auto &sg : segs) {
            sg.key = key;
            if (sg.box >= 0) {
                BoxDesc &B = boxes[sg.box];
                B.key_base = key;
                B.part_off = sg.off;
                key += B.ncx * B.ncy;
                hwm = sg.off + sg.len;
            } else {
                key += 1;
            }
        }
        nkeys = key;
        // every live box small enough for the per-box sort: no global sort
        small_boxes = true;
        for (size_t i = 0; i < boxes.size(); ++i)
            if (box_live[i] && !box_sort_fits(boxes[i].ncx * boxes[i].ncy, boxes[i].n)) small_boxes = false;
        if (nkeys + 1 > cell_cap) {
            int nc = std::max(nkeys + 1, cell_cap * 2);
            dfree(cell_start);
            if (dalloc(&cell_start, nc)) return SPH_ERR_NOMEM;
            cell_cap = nc;
            invalidate_graph();
        }
        if ((int)segs.size() > seg_cap) {
            int nc = std::max((int)segs.size(), seg_cap * 2);
            dfree(d_segs);
            if (dalloc(&d_segs, nc)) return SPH_ERR_NOMEM;
            seg_cap = nc;
            invalidate_graph();
        }
        if (ensure_tables((int)boxes.size())) return SPH_ERR_NOMEM;
        if (!boxes.empty())
            HIPCHK(hipMemcpyAsync(d_boxes, boxes.data(), sizeof(BoxDesc) * boxes.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
        HIPCHK(hipMemcpyAsync(d_segs, segs.data(), sizeof(SegDesc) * segs.size(), hipMemcpyHostToDevice,
                              ctx->stream));
        launch_slot_box(view(), ctx->stream);
        // copies from pageable host memory are staged before hipMemcpyAsync
        // returns, so the tables may change right away; no wait
        // the slot count, key count and segment count are baked into
        // captured launches
        check_graph_sig();
        return SPH_OK;
    }

    // pinned host staging for combined downloads
    int host_stage(size_t bytes)
    {
        if (bytes <= h_stage_bytes) return SPH_OK;
        (void)hipStreamSynchronize(ctx->stream);
        if (h_stage) (void)hipHostFree(h_stage);
        h_stage = nullptr;
        h_stage_bytes = 0;
        if (hipHostMalloc((void **)&h_stage, bytes, hipHostMallocDefault) != hipSuccess) return SPH_ERR_NOMEM;
        h_stage_bytes = bytes;
        return SPH_OK;
    }

    int stage(size_t bytes)
    {
        if (bytes <= stage_bytes) return SPH_OK;
        (void)hipStreamSynchronize(ctx->stream);  // uploads do not wait: let them finish with it
        dfree(d_stage);
        stage_bytes = 0;
        if (dalloc((char **)&d_stage, bytes)) return SPH_ERR_NOMEM;
        stage_bytes = bytes;
        return SPH_OK;
    }

    // ---- timing ----------------------------------------------------------
    int mark(int timer, bool begin)
    {
        if (!ctx->timing) return SPH_OK;
        if (ev_used >= ev_pool.size()) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            ev_pool.push_back(ev);
        }
        size_t idx = ev_used++;
        HIPCHK(hipEventRecord(ev_pool[idx], ctx->stream));
        if (begin) ev_marks.push_back({timer, {idx, (size_t)-1}});
        else if (!ev_marks.empty()) ev_marks.back().second.second = idx;
        return SPH_OK;
    }

    int collect_timing()
    {
        if (!ctx->timing || ev_marks.empty()) {
            ev_used = 0;
            ev_marks.clear();
            return SPH_OK;
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (auto &mk : ev_marks) {
            if (mk.second.second == (size_t)-1) continue;
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev_pool[mk.second.first], ev_pool[mk.second.second]));
            t_acc[mk.first] += ms;
        }
        ev_marks.clear();
        ev_used = 0;
        return SPH_OK;
    }

    // ---- passes ----------------------------------------------------------
    // Cell order and neighbour lists of every box that builds in this launch
    // (see building() in the kernels); the others keep their order.
    void enqueue_build(int force_mask)
    {
        host_dyn_ok = false;  // the build stage writes box states
        hipStream_t st = ctx->stream;
        ArenaView A = view();
        if (hwm == 0) return;
        // the sort runs over the M building slots (M read on the device);
        // the other kernels run over the slots and skip non-building boxes
        launch_build_plan(A, force_mask, st);
        if (small_boxes) {
            launch_box_sort(A, force_mask, (int)boxes.size(), st);
        } else {
            launch_keys(A, force_mask, st);
            radix_sort_pairs(keys_in, k_out, v_out, k_alt, v_alt, hwm, bits_for(nkeys), sort_tmp, st, plan_n + 1);
        }
        launch_permute(A, force_mask, small_boxes ? 0 : 1, st);  // with the cell starts after the radix sort
        launch_build_lists(A, force_mask, fp32() ? 1 : 0, st);
    }

    // One arena step: INIT boxes get u from T0, RUN boxes one leapfrog step
    // (density, gradients, rates, kick/drift); then per-box energy capture
    // and state advance.
    void enqueue_step(bool with_build)
    {
        hipStream_t st = ctx->stream;
        ArenaView A = view();
        const int f = fp32() ? 1 : 0;
        if (with_build) {
            mark(T_SORT, true);
            enqueue_build(0);
            mark(T_SORT, false);
        }
        mark(T_DENS, true);
        launch_density(A, f, MASK_STEP, unroll, st);
        mark(T_DENS, false);
        mark(T_GRAD, true);
        launch_grad(A, f, MASK_RUN, 0, unroll, st);
        mark(T_GRAD, false);
        mark(T_RATES, true);
        launch_rates(A, f, MASK_RUN, rates_unroll(), fuse_step, st);  // with the leapfrog step when fused
        mark(T_RATES, false);
        if (!fuse_step) {
            mark(T_INT, true);
            launch_leapfrog(A, f, st);
            mark(T_INT, false);
        }
        launch_box_step(A, (int)boxes.size(), st);
    }

    int launch_steps(int64_t count, bool with_build)
    {
        host_dyn_ok = false;  // until every step is launched
        const bool graphs = use_graph && !ctx->timing;
        for (int64_t i = 0; i < count;) {
            int lg = 0;  // runs of plain steps: the largest chunk graph that fits
            if (graphs && !with_build)
                while (lg < PLAIN_LOG2 && (int64_t)2 << lg <= count - i) ++lg;
            const int per = 1 << lg;
            const int idx = lg > 0 ? 1 + lg : (with_build ? 1 : 0);
            i += per;
            if (graphs) {
                // single boxes capture at set-up (prepare_graphs); arenas
                // capture here, at the first step after a relayout
                if (int rc = capture_graph(idx)) return rc;
                HIPCHK(hipGraphLaunch(graph[parity][idx], ctx->stream));
                if (per & 1) swap_state();
            } else {
                enqueue_step(with_build);
                swap_state();
            }
        }
        HIPCHK(hipGetLastError());
        host_dyn_ok = count > 0;  // the last launch is a step's k_box_step
        return SPH_OK;
    }

    // Arena stepping: the build stage runs every step and each box builds on
    // its own schedule.
    int step(int64_t nsteps)
    {
        if (hwm == 0) return SPH_OK;
        if (int rc = launch_steps(nsteps, true)) return rc;
        steps_done += (double)nsteps;
        return collect_timing();
    }

    // The schedule of a new run on box b (k = 0, nsteps steps).  With the
    // lists kept, the run continues their schedule from the current step
    // (next build, build step, speed and acceleration bounds, each shifted
    // by the last run's length): a run that follows another builds where
    // the schedule of one longer run would, not at once and again two steps
    // later, as a box with fresh lists does.  plain = the first steps known
    // not to build.
    int run_start(int b, int64_t nsteps, BoxDyn &D, int64_t &plain)
    {
        memset(&D, 0, sizeof(D));
        D.state = BOX_RUN;
        D.k = 0;
        D.nsteps = nsteps;
        D.next_rebuild = 0;
        plain = 0;
        if (sched_kept) {
            if (int rc = read_dyn()) return rc;
            const BoxDyn &P = h_dyn[b];
            D.next_rebuild = P.next_rebuild - P.k;  // <= 0: build at the first step
            D.built_at = P.built_at - P.k;
            D.vlast = P.vlast;
            D.alast = P.alast;
            D.list16 = P.list16;
            plain = D.next_rebuild > 0 ? D.next_rebuild : 0;
        }
        sched_kept = false;  // the run's schedule is on the device until it ends
        return SPH_OK;
    }

    // One running box: steps without the build stage until the box's next
    // scheduled build (read back after each build); the first `plain` steps
    // are known not to build.
    int step_single(int b, int64_t evals, int64_t plain0)
    {
        int64_t done = 0;
        if (plain0 > 0) {
            done = std::min(plain0, evals);
            if (int rc = launch_steps(done, false)) return rc;
        }
        int held = 0;
        // The list review in the step before a scheduled build may postpone
        // it; a step with the build stage then builds nothing, yet its ten
        // launches cost ~45 us.  Once a build step of this run has come up
        // empty, the schedule is read before each build step instead
        // (postponements recur while the particles move slowly, and are
        // rare in a start-up transient, where the extra read would only
        // delay the build).
        bool ask = false, reviewed = plain0 > 0;
        int64_t builds = -1;
        while (done < evals) {
            if (ask && reviewed) {
                if (int rc = read_dyn()) return rc;
                const BoxDyn &R = h_dyn[b];
                if (R.health) break;
                builds = R.builds;
                if (R.k < R.next_rebuild) {  // postponed: more steps without the build stage
                    const int64_t plain = std::min<int64_t>(R.next_rebuild - R.k, evals - done);
                    if (int rc = launch_steps(plain, false)) return rc;
                    done += plain;
                    continue;
                }
            }
            const int64_t rel = released;
            if (int rc = launch_steps(1, true)) return rc;
            if (int rc = read_dyn()) return rc;
            if (released != rel) {  // the build held the box; it builds again in 32-bit entries
                if (++held > 2) return fail(SPH_ERR_HIP, "list build held repeatedly");
                reviewed = false;
                continue;
            }
            ++done;
            const BoxDyn &D = h_dyn[b];
            if (builds >= 0 && D.builds == builds) ask = true;  // that build step found its build postponed
            builds = D.builds;
            if (done >= evals) break;
            if (D.health) break;
            int64_t plain = D.next_rebuild - D.k;
            if (plain < 0) plain = 0;
            if (plain > evals - done) plain = evals - done;
            if (int rc = launch_steps(plain, false)) return rc;
            done += plain;
            reviewed = plain > 0;
        }
        steps_done += (double)evals;
        return collect_timing();
    }

    // Advance box b's open run by nsteps steps (one rate evaluation each),
    // starting the run (k = 0: half kick) if none is open.  The launches are
    // those sph_step makes for the same steps: plain steps until the box's
    // scheduled build, then a step with the build stage and one read-back of
    // the schedule, so k steps in any number of calls give the bits of one
    // run of k steps.
    int advance(int b, int64_t nsteps)
    {
        if (!open_run) {
            BoxDyn D;
            int64_t plain = 0;
            if (int rc = run_start(b, OPEN_STEPS, D, plain)) return rc;
            if (int rc = write_dyn(b, D)) return rc;
            open_run = true;
            open_plain = plain;
        }
        int64_t done = 0;
        while (done < nsteps) {
            if (open_plain < 0) {
                // the schedule after the last step with the build stage, read
                // when it is needed (at this call, not right after that
                // launch, so the host's work between calls overlaps the step)
                if (int rc = retake_held()) return rc;
                const BoxDyn &D = h_dyn[b];
                if (D.health) {
                    open_run = false;
                    return SPH_OK;  // the caller's health check reports it
                }
                open_plain = D.next_rebuild > D.k ? D.next_rebuild - D.k : 0;
            }
            if (open_plain == 0) {
                if (int rc = launch_steps(1, true)) return rc;
                open_plain = -1;
                ++done;
            } else {
                const int64_t p = std::min(open_plain, nsteps - done);
                if (int rc = launch_steps(p, false)) return rc;
                open_plain -= p;
                done += p;
            }
        }
        steps_done += (double)nsteps;
        stepped = true;
        return collect_timing();
    }

    // Close box b's open run: the rate evaluation at the current step with
    // the closing half kick, so v and u are synchronised with r again.
    int close_run(int b)
    {
        if (!open_run) return SPH_OK;
        open_run = false;
        if (int rc = retake_held()) return rc;  // also waits for the run's steps
        BoxDyn D = h_dyn[b];
        if (D.health || D.state != BOX_RUN) return SPH_OK;
        D.nsteps = D.k;
        if (int rc = write_dyn(b, D)) return rc;
        for (int held = 0;;) {
            const int64_t rel = released;
            if (int rc = launch_steps(1, true)) return rc;
            if (int rc = read_dyn()) return rc;
            if (released == rel) break;  // else a held build: again, in 32-bit entries
            if (++held > 2) return fail(SPH_ERR_HIP, "list build held repeatedly");
        }
        return collect_timing();
    }

    // Read the box states; while the last launched step was a build that
    // held its box (released by the read, now in 32-bit entries), launch
    // that step again, so the steps a caller asked for all take effect.
    int retake_held()
    {
        for (int held = 0;;) {
            const int64_t rel = released;
            if (int rc = read_dyn()) return rc;
            if (released == rel) return SPH_OK;
            if (++held > 2) return fail(SPH_ERR_HIP, "list build held repeatedly");
            if (int rc = launch_steps(1, true)) return rc;
        }
    }

    // The box states, read back (waits for the stream); boxes held by a build
    // are released on the way (release_held).  After a step (host_dyn_ok)
    // the step's k_box_step has left them in host memory and the wait is
    // the whole read; otherwise a device-to-host copy.
    int read_dyn()
    {
        if (boxes.empty()) return SPH_OK;
        if (h_stepped && host_dyn_ok) {
            HIPCHK(hipStreamSynchronize(ctx->stream));
            memcpy(h_dyn, h_stepped, sizeof(BoxDyn) * boxes.size());
        } else {
            HIPCHK(hipMemcpyAsync(h_dyn, d_dyn, sizeof(BoxDyn) * boxes.size(), hipMemcpyDeviceToHost,
                                  ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
        }
        return release_held();
    }

    // A box whose build met a neighbour offset beyond 16 bits held that step
    // (BoxDyn::relist): switch it to 32-bit entries and clear the flag; its
    // next step builds again, with the same entries in the same order.  The
    // copies are stream-ordered before any later launch.
    int release_held()
    {
        static const int zero = 0;
        for (size_t b = 0; b < boxes.size(); ++b) {
            if (!box_live[b] || !h_dyn[b].relist) continue;
            host_dyn_ok = false;
            boxes[b].build16 = 0;
            HIPCHK(hipMemcpyAsync(d_boxes + b, &boxes[b], sizeof(BoxDesc), hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipMemcpyAsync((char *)(d_dyn + b) + offsetof(BoxDyn, relist), &zero, sizeof(int),
                                  hipMemcpyHostToDevice, ctx->stream));
            h_dyn[b].relist = 0;
            ++released;
        }
        return SPH_OK;
    }

    // Build stage outside a run (the stand-alone evaluations and the pair
    // export) for the boxes in force_mask; a held build is repeated in
    // 32-bit entries, so the lists are whole when this returns.
    int build_now(int force_mask)
    {
        for (int held = 0;;) {
            const int64_t rel = released;
            enqueue_build(force_mask);
            HIPCHK(hipGetLastError());
            if (int rc = read_dyn()) return rc;
            if (released == rel) return SPH_OK;
            if (++held > 2) return fail(SPH_ERR_HIP, "list build held repeatedly");
        }
    }

    // h_dyn is pinned and copies run in stream order (a later read_dyn lands
    // after this copy), so no wait is needed here
    int write_dyn(int b, const BoxDyn &D)
    {
        h_dyn[b] = D;
        host_dyn_ok = false;
        HIPCHK(hipMemcpyAsync(d_dyn + b, h_dyn + b, sizeof(BoxDyn), hipMemcpyHostToDevice, ctx->stream));
        return SPH_OK;
    }

    // ---- boxes -------------------------------------------------------------
    int add_box(int64_t n, const sph_params *prm, int *box_out);
    int remove_box(int b);
    int upload(int b, int field, const double *host, int ncomp);
    int download(int b, int field, double *host, int ncomp);
};

int fill_desc(const sph_params *p, int64_t n, BoxDesc &B)
{
    if (!(p->Lx > 0) || !(p->Ly > 0) || !(p->h > 0) || !(p->hl > 0) || !(p->dt >= 0) || !(p->kB > 0) ||
        !(p->rc >= 0) || !(p->tol >= 0) || !std::isfinite(p->Lx) || !std::isfinite(p->Ly))
        return fail(SPH_ERR_ARG, "invalid sph_params: need Lx, Ly, h, hl, kB > 0, rc, tol, dt >= 0");
    if (p->precision != SPH_FP64 && p->precision != SPH_FP32C)
        return fail(SPH_ERR_ARG, "invalid precision");
    if (!(p->skin >= 0) || !std::isfinite(p->skin)) return fail(SPH_ERR_ARG, "invalid skin (need >= 0)");
    memset(&B, 0, sizeof(B));
    const double PI = 3.14159265358979323846;
    B.Lx = p->Lx; B.Ly = p->Ly; B.hLx = p->Lx / 2.; B.hLy = p->Ly / 2.;
    B.R2list = p->rc * p->rc + p->tol * p->tol;  // neighbour_list.py:156-157,178
    B.h2 = p->h * p->h; B.hl2 = p->hl * p->hl;
    B.cut2 = std::min(B.R2list, std::max(B.h2, B.hl2));
    const double rcut = std::sqrt(B.cut2);
    B.skin = p->skin;
    B.rnl2 = (rcut + p->skin) * (rcut + p->skin);
    B.hsk2 = (p->h + p->skin) * (p->h + p->skin);
    B.hskin2 = 0.25 * p->skin * p->skin;
    // listed pairs are closer than r_cut + skin at the build and each end
    // moves at most skin / 2 before the next, so within r_cut + 2 skin
    B.rin = 2.0 * (rcut + p->skin);
    // FP32C boxes up to the box-relative limit per side mirror positions as
    // box-relative fixed point (2^32 units per side)
    B.lolong = (p->precision == SPH_FP32C && (p->Lx > FP32C_HI_ONLY_MAX_L || p->Ly > FP32C_HI_ONLY_MAX_L)) ? 1 : 0;
    // cells cover both the pair-export radius and the list radius
    const double Rc = std::max(std::sqrt(B.R2list), rcut + p->skin) * (1.0 + 1e-9);
    int ncx = (int)std::floor(p->Lx / Rc), ncy = (int)std::floor(p->Ly / Rc);
    if (Rc <= 0) { ncx = 1; ncy = 1; }
    if (ncx < 3) ncx = 1;
    if (ncy < 3) ncy = 1;
    if ((long long)ncx * ncy > (1ll << 26))
        return fail(SPH_ERR_ARG, "box has too many cells for the cell list (> 2^26)");
    if (p->precision == SPH_FP32C && (ncx > 65535 || ncy > 65535))
        return fail(SPH_ERR_ARG, "FP32C boxes hold at most 65535 cells per axis");
    B.ncx = ncx; B.ncy = ncy; B.sx = ncx >= 3 ? 1 : 0; B.sy = ncy >= 3 ? 1 : 0;
    B.cwx = p->Lx / ncx; B.cwy = p->Ly / ncy; B.icwx = ncx / p->Lx; B.icwy = ncy / p->Ly;
    B.h = p->h; B.hl = p->hl; B.ih = 1.0 / p->h; B.ihl = 1.0 / p->hl;
    B.qh = 5.0 / (PI * p->h * p->h);        // spkernel.py:98
    B.qhl = 5.0 / (PI * p->hl * p->hl);
    B.fh = -12.0 * B.qh / (p->h * p->h);    // spkernel.py:112-113, factored
    B.fhl = -12.0 * B.qhl / (p->hl * p->hl);
    B.a = p->a; B.b = p->b; B.kB = p->kB; B.eta = p->eta; B.zeta = p->zeta; B.kcond = p->kcond;
    B.dt = p->dt;
    B.n = (int)n;
    B.terms = (int)(p->terms & SPH_ALL_TERMS);
    B.precision = p->precision;
    return SPH_OK;
}

// Neighbour-list entries per slot for a box: 2.5x the mean count at the
// box's number density plus 32, rounded up to 32 (at least 64).
int list_entries_for(const sph_params *p, int64_t n)
{
    const double rcut = std::sqrt(std::min(p->rc * p->rc + p->tol * p->tol, std::max(p->h * p->h, p->hl * p->hl)));
    const double rnl = rcut + p->skin;
    const double dens = (double)n / (p->Lx * p->Ly);
    const double est = 3.14159265358979323846 * rnl * rnl * dens;
    int l = (int)std::ceil((2.5 * est + 32.0) / 32.0) * 32;
    return std::max(64, std::min(l, 4096));
}

int Arena::add_box(int64_t n, const sph_params *prm, int *box_out)
{
    if (n <= 0 || n > cap) return fail(SPH_ERR_ARG, "box size out of range");
    BoxDesc B;
    if (int rc = fill_desc(prm, n, B)) return rc;
    if (prm->precision != precision) return fail(SPH_ERR_ARG, "box precision differs from the arena's");
    B.build16 = list_bits == 16 ? 1 : 0;
    {
        // mean neighbours within r_cut + skin at the box's number density:
        // a box whose mean alone takes over half the list entries is refused
        // here (a liquid region at twice the mean density would overflow
        // them) instead of failing its first build with HEALTH_LISTCAP
        const double mean = 3.14159265358979323846 * B.rnl2 * (double)n / (prm->Lx * prm->Ly);
        if (mean > 0.5 * lcap)
            return fail(SPH_ERR_ARG, "box too dense for the neighbour lists: " + std::to_string(mean) +
                                         " mean neighbours within r_cut + skin against " + std::to_string(lcap) +
                                         " list entries per particle");
    }
    // first fit over free segments
    int si = -1;
    for (size_t i = 0; i < segs.size(); ++i)
        if (segs[i].box < 0 && segs[i].len >= n) { si = (int)i; break; }
    if (si < 0) return fail(SPH_ERR_NOMEM, "arena has no free slot range of " + std::to_string(n));
    int b = -1;
    for (size_t i = 0; i < box_live.size(); ++i)
        if (!box_live[i]) { b = (int)i; break; }
    if (b < 0) {
        b = (int)boxes.size();
        boxes.push_back(B);
        box_live.push_back(1);
    } else {
        boxes[b] = B;
        box_live[b] = 1;
    }
    SegDesc fs = segs[si];
    SegDesc bs{fs.off, (int)n, b, 0};
    std::vector<SegDesc> ns(segs.begin(), segs.begin() + si);
    ns.push_back(bs);
    if (fs.len > n) ns.push_back(SegDesc{fs.off + (int)n, fs.len - (int)n, -1, 0});
    ns.insert(ns.end(), segs.begin() + si + 1, segs.end());
    segs.swap(ns);
    if (int rc = relayout()) return rc;
    BoxDyn D;
    memset(&D, 0, sizeof(D));
    D.state = BOX_DONE;
    if (int rc = write_dyn(b, D)) return rc;
    launch_init_slots(view(), boxes[b].part_off, (int)n, ctx->stream);
    HIPCHK(hipGetLastError());
    *box_out = b;
    return SPH_OK;
}

int Arena::remove_box(int b)
{
    if (b < 0 || b >= (int)boxes.size() || !box_live[b]) return fail(SPH_ERR_ARG, "no such box");
    for (auto &sg : segs)
        if (sg.box == b) sg.box = -1;
    // merge adjacent free segments
    std::vector<SegDesc> ns;
    for (auto &sg : segs) {
        if (!ns.empty() && ns.back().box < 0 && sg.box < 0) ns.back().len += sg.len;
        else ns.push_back(sg);
    }
    segs.swap(ns);
    box_live[b] = 0;
    boxes[b].n = 0;
    BoxDyn D;
    memset(&D, 0, sizeof(D));
    D.state = BOX_FREE;
    if (int rc = write_dyn(b, D)) return rc;
    return relayout();
}

int Arena::upload(int b, int field, const double *host, int ncomp)
{
    const int n = boxes[b].n;
    sched_kept = false;  // new positions or new speed and acceleration bounds
    const size_t bytes = sizeof(double) * (size_t)n * ncomp;
    if (stage(bytes)) return SPH_ERR_NOMEM;
    HIPCHK(hipMemcpyAsync(d_stage, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    launch_upload(view(), boxes[b].part_off, n, field, d_stage, ctx->stream);
    HIPCHK(hipGetLastError());
    if (field == 2) {
        // equal masses (every configuration of the reference's drivers):
        // the density pass then skips the per-neighbour mass fetch
        bool uni = n > 0 && host[0] > 0.0 && std::isfinite(host[0]);
        for (int i = 1; uni && i < n; ++i) uni = host[i] == host[0];
        boxes[b].munif = uni ? host[0] : 0.0;
        HIPCHK(hipMemcpyAsync(d_boxes + b, &boxes[b], sizeof(BoxDesc), hipMemcpyHostToDevice, ctx->stream));
    }
    // the pageable source is staged before the copy call returns and the
    // staging buffer is reused in stream order: no wait
    return SPH_OK;
}

int Arena::download(int b, int field, double *host, int ncomp)
{
    const int n = boxes[b].n;
    const size_t bytes = sizeof(double) * (size_t)n * ncomp;
    if (stage(bytes)) return SPH_ERR_NOMEM;
    launch_download(view(), b, n, field, d_stage, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, d_stage, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int check_health(Arena &A, int b)
{
    if (int rc = A.read_dyn()) return rc;
    const int h = A.h_dyn[b].health;
    if (h == 0) return SPH_OK;
    std::string msg = "numerical failure in box " + std::to_string(b) + ":";
    if (h & HEALTH_NONFINITE) msg += " non-finite density/rates/state;";
    if (h & HEALTH_EOS_POLE) msg += " vdW pole rho*b >= 1 reached;";
    if (h & HEALTH_SKIN) msg += " a particle moved more than half the neighbour-list skin between builds;";
    if (h & HEALTH_LISTCAP) msg += " neighbour-list capacity exceeded (density too high for the list);";
    return fail(SPH_ERR_NUMERIC, msg);
}

}  // namespace

struct sph_box {
    Arena ar;
    int b = 0;
    sph_params prm;
};

struct sph_arena {
    Arena ar;
};

namespace {
// Close a single box's open run (sph_box_advance) before a call that reads
// or replaces its state: v and u are then synchronised with r again.
int settle(sph_box *box)
{
    Arena &A = box->ar;
    if (!A.open_run) return SPH_OK;
    HIPCHK(hipSetDevice(A.ctx->device));
    if (int rc = A.close_run(box->b)) return rc;
    if (int rc = check_health(A, box->b)) return rc;
    A.sched_kept = true;
    return SPH_OK;
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

const char *sph_last_error(void) { return g_err.c_str(); }

const char *sph_version(void) { return "pyticles_hip 0.1 (gfx950)"; }

int sph_ctx_create(int device, sph_ctx **out)
{
    if (!out) return fail(SPH_ERR_ARG, "null out pointer");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(SPH_ERR_ARG, "device " + std::to_string(device) + " not present (" + std::to_string(ndev) + " devices)");
    HIPCHK(hipSetDevice(device));
    if (int rc = noalloc_scratch_setup(device)) return rc;
    sph_ctx *c = new sph_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(SPH_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return SPH_OK;
}

int sph_ctx_destroy(sph_ctx *ctx)
{
    if (!ctx) return SPH_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    dfree(ctx->d_scratch);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SPH_OK;
}

int sph_ctx_set_timing(sph_ctx *ctx, int on)
{
    if (!ctx) return fail(SPH_ERR_ARG, "null context");
    ctx->timing = on ? 1 : 0;
    return SPH_OK;
}

// ---- one box ---------------------------------------------------------------

int sph_box_create(sph_ctx *ctx, int64_t n, const sph_params *prm, sph_box **out)
{
    if (!ctx || !prm || !out) return fail(SPH_ERR_ARG, "null argument");
    *out = nullptr;
    if (n <= 0 || n > (1ll << 30)) return fail(SPH_ERR_ARG, "n out of range");
    HIPCHK(hipSetDevice(ctx->device));
    sph_box *bx = new sph_box();
    bx->ar.ctx = ctx;
    bx->ar.keep_host_dyn = true;
    bx->prm = *prm;
    int rc = bx->ar.alloc((int)n, prm->precision, list_entries_for(prm, n));
    if (!rc) rc = bx->ar.add_box(n, prm, &bx->b);
    if (!rc) rc = bx->ar.prepare_graphs();  // no capture inside a later sph_step
    if (rc) {
        std::string keep = g_err;
        bx->ar.release();
        delete bx;
        g_err = keep;
        return rc;
    }
    *out = bx;
    return SPH_OK;
}

int sph_box_destroy(sph_box *box)
{
    if (!box) return SPH_OK;
    (void)hipSetDevice(box->ar.ctx->device);
    (void)hipStreamSynchronize(box->ar.ctx->stream);
    box->ar.release();
    delete box;
    return SPH_OK;
}

int sph_box_set_params(sph_box *box, const sph_params *prm)
{
    if (!box || !prm) return fail(SPH_ERR_ARG, "null argument");
    if (int rc = settle(box)) return rc;
    if (prm->precision != box->ar.precision) return fail(SPH_ERR_ARG, "precision cannot change");
    BoxDesc B;
    if (int rc = fill_desc(prm, box->ar.boxes[box->b].n, B)) return rc;
    B.munif = box->ar.boxes[box->b].munif;  // state, not parameters
    B.T0 = box->ar.boxes[box->b].T0;
    B.build16 = box->ar.boxes[box->b].build16;
    box->ar.boxes[box->b] = B;
    box->prm = *prm;
    if (int rc = box->ar.relayout()) return rc;
    return box->ar.prepare_graphs();
}

static int check_positions(const sph_box *box, const double *r)
{
    const int n = box->ar.boxes[box->b].n;
    const double Lx = box->prm.Lx, Ly = box->prm.Ly;
    for (int i = 0; i < n; ++i) {
        const double x = r[2 * i], y = r[2 * i + 1];
        if (!(x >= 0.0 && x < Lx && y >= 0.0 && y < Ly))
            return fail(SPH_ERR_ARG, "position of particle " + std::to_string(i) +
                                         " lies outside [0, L); wrap positions into the periodic box first");
    }
    return SPH_OK;
}

int sph_box_upload(sph_box *box, const double *r, const double *v, const double *m, const double *u)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    if (int rc = settle(box)) return rc;
    HIPCHK(hipSetDevice(box->ar.ctx->device));
    if (r) {
        if (int rc = check_positions(box, r)) return rc;
        if (int rc = box->ar.upload(box->b, 0, r, 2)) return rc;
    }
    if (v) if (int rc = box->ar.upload(box->b, 1, v, 2)) return rc;
    if (m) if (int rc = box->ar.upload(box->b, 2, m, 1)) return rc;
    if (u) if (int rc = box->ar.upload(box->b, 3, u, 1)) return rc;
    return SPH_OK;
}

int sph_box_download(sph_box *box, double *r, double *v, double *u, double *rho, double *rho_lr,
                     double *T, double *p, double *pco, double *acc, double *udot, double *gradv,
                     double *jq, double *P)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    // positions are at the current integer time in an open run too; every
    // other field needs the run closed first
    const bool only_r = !v && !u && !rho && !rho_lr && !T && !p && !pco && !acc && !udot && !gradv && !jq && !P;
    if (!only_r)
        if (int rc = settle(box)) return rc;
    HIPCHK(hipSetDevice(box->ar.ctx->device));
    struct F { double *dst; int field, nc; } fs[] = {
        {r, 0, 2}, {v, 1, 2}, {u, 3, 1}, {rho, 4, 1}, {rho_lr, 5, 1}, {T, 6, 1}, {p, 7, 1},
        {pco, 8, 1}, {acc, 9, 2}, {udot, 10, 1}, {gradv, 11, 4}, {jq, 12, 2}, {P, 13, 4}};
    for (auto &f : fs)
        if (f.dst)
            if (int rc = box->ar.download(box->b, f.field, f.dst, f.nc)) return rc;
    return SPH_OK;
}

// ---- (a) neighbour list ---------------------------------------------------

int sph_nlist_build(sph_box *box, int64_t cap, int64_t *iap, int64_t *nip_out)
{
    if (!box || !nip_out || cap < 0 || (cap > 0 && !iap)) return fail(SPH_ERR_ARG, "invalid argument");
    if (int rc = settle(box)) return rc;
    Arena &A = box->ar;
    hipStream_t st = A.ctx->stream;
    HIPCHK(hipSetDevice(A.ctx->device));
    const int n = A.boxes[box->b].n;
    A.sched_kept = false;  // lists rebuilt outside a run's schedule
    A.mark(T_PAIRS, true);
    if (int rc = A.build_now(MASK_ALL)) return rc;
    if (!A.d_inv) {
        if (dalloc(&A.d_inv, A.cap) || dalloc(&A.d_cnt, A.cap) || dalloc(&A.d_off, A.cap) ||
            dalloc(&A.d_grand, 1))
            return SPH_ERR_NOMEM;
        A.scan_tmp_bytes = exscan_temp_bytes(A.cap);
        if (dalloc((char **)&A.scan_tmp, A.scan_tmp_bytes)) return SPH_ERR_NOMEM;
    }
    ArenaView V = A.view();
    launch_inverse(A.id_s, A.boxes[box->b].part_off, n, A.d_inv, st);
    launch_pairs(V, box->b, n, A.d_inv, A.d_cnt, nullptr, nullptr, 0, st);
    exclusive_scan_u32(A.d_cnt, A.d_off, n, A.d_grand, A.scan_tmp, st);
    unsigned grand = 0;
    HIPCHK(hipMemcpyAsync(&grand, A.d_grand, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int64_t nip = grand;
    if ((size_t)nip > A.jbuf_cap) {
        dfree(A.d_jbuf);
        A.jbuf_cap = 0;
        if (dalloc(&A.d_jbuf, (size_t)nip + 1)) return SPH_ERR_NOMEM;
        A.jbuf_cap = (size_t)nip + 1;
    }
    const int64_t nout = std::min<int64_t>(nip, cap);
    if ((size_t)nout > A.iap_cap) {
        dfree(A.d_iap);
        A.iap_cap = 0;
        if (dalloc(&A.d_iap, 2 * (size_t)nout + 2)) return SPH_ERR_NOMEM;
        A.iap_cap = (size_t)nout;
    }
    launch_pairs(V, box->b, n, A.d_inv, A.d_cnt, A.d_off, A.d_jbuf, 1, st);
    launch_pairs_emit(n, A.d_off, A.d_cnt, A.d_jbuf, A.d_iap, (long long)nout, st);
    HIPCHK(hipGetLastError());
    if (nout > 0)
        HIPCHK(hipMemcpyAsync(iap, A.d_iap, sizeof(long long) * 2 * nout, hipMemcpyDeviceToHost, st));
    A.mark(T_PAIRS, false);
    HIPCHK(hipStreamSynchronize(st));
    if (int rc = A.collect_timing()) return rc;
    *nip_out = nip;
    if (nip > cap)
        return fail(SPH_ERR_CAP, "pair capacity " + std::to_string(cap) + " < " + std::to_string(nip) + " pairs");
    return SPH_OK;
}

int sph_separations(sph_box *box, int64_t nip, const int64_t *iap, double *drij, double *rij,
                    double *rsq, double *dv)
{
    if (!box || nip < 0 || (nip > 0 && !iap)) return fail(SPH_ERR_ARG, "invalid argument");
    if (int rc = settle(box)) return rc;
    if (nip == 0) return SPH_OK;
    Arena &A = box->ar;
    hipStream_t st = A.ctx->stream;
    HIPCHK(hipSetDevice(A.ctx->device));
    const int n = A.boxes[box->b].n;
    for (int64_t k = 0; k < 2 * nip; ++k)
        if (iap[k] < 0 || iap[k] >= n) return fail(SPH_ERR_ARG, "pair index out of range");
    if (!A.d_inv && dalloc(&A.d_inv, A.cap)) return SPH_ERR_NOMEM;
    // device layout: iap (2*nip int64) then up to 6*nip doubles of outputs
    const size_t bytes = sizeof(long long) * 2 * nip + sizeof(double) * 6 * nip;
    if (A.stage(bytes)) return SPH_ERR_NOMEM;
    long long *d_iap = (long long *)A.d_stage;
    double *o = (double *)(d_iap + 2 * nip);
    double *d_drij = drij ? o : nullptr;
    double *d_dv = dv ? o + 2 * nip : nullptr;
    double *d_rij = rij ? o + 4 * nip : nullptr;
    double *d_rsq = rsq ? o + 5 * nip : nullptr;
    HIPCHK(hipMemcpyAsync(d_iap, iap, sizeof(long long) * 2 * nip, hipMemcpyHostToDevice, st));
    launch_inverse(A.id, A.boxes[box->b].part_off, n, A.d_inv, st);
    launch_separations(A.view(), box->b, A.d_inv, nip, d_iap, d_drij, d_rij, d_rsq, d_dv, st);
    HIPCHK(hipGetLastError());
    if (drij) HIPCHK(hipMemcpyAsync(drij, d_drij, sizeof(double) * 2 * nip, hipMemcpyDeviceToHost, st));
    if (dv) HIPCHK(hipMemcpyAsync(dv, d_dv, sizeof(double) * 2 * nip, hipMemcpyDeviceToHost, st));
    if (rij) HIPCHK(hipMemcpyAsync(rij, d_rij, sizeof(double) * nip, hipMemcpyDeviceToHost, st));
    if (rsq) HIPCHK(hipMemcpyAsync(rsq, d_rsq, sizeof(double) * nip, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return SPH_OK;
}

void lucy_kernel_3d(float r, const float *dx, float h, float *w, float *dwdx)
{
    const float q = 105.0f / (3.14159265358979323846f * 16.0f * (h * h * h));
    r = std::fabs(r);
    if (r > h) {
        *w = 0.0f;
        dwdx[0] = dwdx[1] = dwdx[2] = 0.0f;
        return;
    }
    const float t = 1.0f - r / h;
    *w = q * (1.0f + 3.0f * r / h) * (t * t * t);
    if (r == 0.0f) {
        dwdx[0] = dwdx[1] = dwdx[2] = 0.0f;
        return;
    }
    // dW/dr = q (-12 r^3 / h^4 + 24 r^2 / h^3 - 12 r / h^2)
    const float dwdr = q * ((-12.0f / (h * h * h * h)) * (r * r * r) + (24.0f / (h * h * h)) * (r * r) -
                            12.0f * r / (h * h));
    for (int c = 0; c < 3; ++c) dwdx[c] = dwdr * dx[c] / r;
}

int sph_fp32_accumulate(sph_ctx *ctx, int64_t n, const double *x, int unroll, double *kahan, double *plain)
{
    if (!ctx || n < 0 || n > (1ll << 24) || (n > 0 && !x) || !kahan || !plain || unroll < 1 || unroll > 64)
        return fail(SPH_ERR_ARG, "invalid argument (n <= 2^24 terms, unroll 1..64)");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double) * ((size_t)n + 2);
    if (ctx->scratch(bytes)) return SPH_ERR_NOMEM;
    double *d_x = ctx->d_scratch, *d_out = ctx->d_scratch + n;
    if (n > 0) HIPCHK(hipMemcpyAsync(d_x, x, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    launch_fp32_accumulate(n, d_x, unroll, d_out, ctx->stream);
    HIPCHK(hipGetLastError());
    double out[2];
    HIPCHK(hipMemcpyAsync(out, d_out, sizeof(out), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *kahan = out[0];
    *plain = out[1];
    return SPH_OK;
}

int sph_minimum_image(sph_ctx *ctx, int64_t np, double *dr, double Lx, double Ly)
{
    if (!ctx || np < 0 || (np > 0 && !dr)) return fail(SPH_ERR_ARG, "invalid argument");
    if (np == 0) return SPH_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double) * 2 * np;
    if (ctx->scratch(bytes)) return SPH_ERR_NOMEM;
    HIPCHK(hipMemcpyAsync(ctx->d_scratch, dr, bytes, hipMemcpyHostToDevice, ctx->stream));
    launch_min_image(np, ctx->d_scratch, Lx, Ly, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dr, ctx->d_scratch, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

// ---- (b) kernel, density, EOS ----------------------------------------------

int sph_kernel_lucy(sph_ctx *ctx, int64_t np, const double *r, const double *dx, double h, int dim,
                    double *w, double *dwdx)
{
    if (!ctx || np < 0 || dim < 1 || dim > 3 || !(h > 0) || (np > 0 && (!r || !dx || !w || !dwdx)))
        return fail(SPH_ERR_ARG, "invalid argument (dim must be 1..3, h > 0)");
    if (np == 0) return SPH_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t nb = sizeof(double) * np;
    if (ctx->scratch(nb * (2 + 2 * dim))) return SPH_ERR_NOMEM;
    double *d_r = ctx->d_scratch, *d_dx = d_r + np, *d_w = d_dx + np * dim, *d_g = d_w + np;
    HIPCHK(hipMemcpyAsync(d_r, r, nb, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(d_dx, dx, nb * dim, hipMemcpyHostToDevice, ctx->stream));
    launch_lucy(np, d_r, d_dx, h, dim, d_w, d_g, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(w, d_w, nb, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(dwdx, d_g, nb * dim, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

// Build (cell order + lists) then density, optional gradients and rates of
// the current state; every derived field ends in the new slot order.
static int eval_box(sph_box *box, int grad, int rates)
{
    if (int rc = settle(box)) return rc;
    Arena &A = box->ar;
    hipStream_t st = A.ctx->stream;
    HIPCHK(hipSetDevice(A.ctx->device));
    const int f = A.fp32() ? 1 : 0;
    ArenaView V = A.view();
    A.sched_kept = false;  // lists rebuilt outside a run's schedule
    A.host_dyn_ok = false;  // the passes below may flag box health
    A.mark(T_SORT, true);
    if (int rc = A.build_now(MASK_ALL)) return rc;
    A.mark(T_SORT, false);
    A.mark(T_DENS, true);
    launch_density(V, f, MASK_ALL, A.unroll, st);
    A.mark(T_DENS, false);
    if (grad) {
        A.mark(T_GRAD, true);
        launch_grad(V, f, MASK_ALL, 1, A.unroll, st);
        A.mark(T_GRAD, false);
    }
    if (rates) {
        A.mark(T_RATES, true);
        launch_rates(V, f, MASK_ALL, A.rates_unroll(), 0, st);
        A.mark(T_RATES, false);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    if (int rc = A.collect_timing()) return rc;
    return check_health(A, box->b);
}

int sph_density(sph_box *box)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    return eval_box(box, 0, 0);
}

int sph_gradients(sph_box *box)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    return eval_box(box, 1, 0);
}

int sph_rates(sph_box *box, unsigned terms)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    Arena &A = box->ar;
    BoxDesc &B = A.boxes[box->b];
    const int keep = B.terms;
    B.terms = (int)(terms & SPH_ALL_TERMS);
    if (int rc = A.relayout()) return rc;
    int rc = eval_box(box, (terms & (SPH_VISCOSITY | SPH_CONDUCTION)) ? 1 : 0, 1);
    std::string msg = g_err;
    B.terms = keep;
    int rc2 = A.relayout();
    if (rc) {
        g_err = msg;
        return rc;
    }
    return rc2;
}

int sph_init_thermal(sph_box *box, const double *T, double T0)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    Arena &A = box->ar;
    const int n = A.boxes[box->b].n;
    if (int rc = eval_box(box, 0, 0)) return rc;
    double *dT = nullptr;
    if (T) {
        if (A.stage(sizeof(double) * n)) return SPH_ERR_NOMEM;
        HIPCHK(hipMemcpyAsync(A.d_stage, T, sizeof(double) * n, hipMemcpyHostToDevice, A.ctx->stream));
        dT = A.d_stage;
    }
    launch_set_u(A.view(), box->b, n, dT, T0, A.ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(A.ctx->stream));
    return SPH_OK;
}

// ---- time stepping -----------------------------------------------------------

int sph_step(sph_box *box, int64_t nsteps)
{
    if (!box || nsteps < 0) return fail(SPH_ERR_ARG, "invalid argument");
    if (nsteps == 0) return SPH_OK;
    Arena &A = box->ar;
    HIPCHK(hipSetDevice(A.ctx->device));
    if (A.open_run) {  // continue the open run by nsteps, then close it
        if (int rc = A.advance(box->b, nsteps)) return rc;
        if (int rc = A.close_run(box->b)) return rc;
    } else {
        BoxDyn D;
        int64_t plain = 0;
        if (int rc = A.run_start(box->b, nsteps, D, plain)) return rc;
        if (int rc = A.write_dyn(box->b, D)) return rc;
        // nsteps leapfrog steps take nsteps + 1 rate evaluations
        if (int rc = A.step_single(box->b, nsteps + 1, plain)) return rc;
        A.steps_done -= 1.0;
        A.stepped = true;
    }
    if (int rc = check_health(A, box->b)) return rc;
    A.sched_kept = true;
    return SPH_OK;
}

int sph_box_advance(sph_box *box, int64_t nsteps)
{
    if (!box || nsteps < 0) return fail(SPH_ERR_ARG, "invalid argument");
    if (nsteps == 0) return SPH_OK;
    Arena &A = box->ar;
    HIPCHK(hipSetDevice(A.ctx->device));
    if (int rc = A.advance(box->b, nsteps)) return rc;
    if (!A.open_run) return check_health(A, box->b);  // a health bit ended the run
    return SPH_OK;
}

int sph_box_sync(sph_box *box)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    return settle(box);
}

int sph_box_run_stats(sph_box *box, int64_t *steps, int64_t *builds, int64_t *extends, int *open)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    Arena &A = box->ar;
    HIPCHK(hipSetDevice(A.ctx->device));
    if (int rc = A.read_dyn()) return rc;
    const BoxDyn &D = A.h_dyn[box->b];
    if (steps) *steps = D.k;
    if (builds) *builds = D.builds;
    if (extends) *extends = D.extends;
    if (open) *open = A.open_run ? 1 : 0;
    return SPH_OK;
}

int sph_energy(sph_box *box, double *ke, double *ue, double *etot)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    if (int rc = settle(box)) return rc;
    Arena &A = box->ar;
    HIPCHK(hipSetDevice(A.ctx->device));
    double out[3];
    launch_energy_now(A.view(), box->b, A.d_e3, A.ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, A.d_e3, sizeof(out), hipMemcpyDeviceToHost, A.ctx->stream));
    HIPCHK(hipStreamSynchronize(A.ctx->stream));
    if (ke) *ke = out[0];
    if (ue) *ue = out[1];
    if (etot) *etot = out[2];
    return SPH_OK;
}

static int read_timing(Arena &A, double *out, int n, int reset)
{
    if (n < 0 || n > 7 || (n > 0 && !out)) return fail(SPH_ERR_ARG, "invalid argument");
    const double vals[7] = {A.t_acc[T_SORT], A.t_acc[T_DENS], A.t_acc[T_GRAD], A.t_acc[T_RATES],
                            A.t_acc[T_PAIRS], A.steps_done, A.t_acc[T_INT]};
    for (int i = 0; i < n; ++i) out[i] = vals[i];
    if (reset) {
        for (double &t : A.t_acc) t = 0;
        A.steps_done = 0;
    }
    return SPH_OK;
}

int sph_box_timing(sph_box *box, double *out, int n, int reset)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    return read_timing(box->ar, out, n, reset);
}

int sph_box_time_pass(sph_box *box, int pass, int reps, double *us_per_launch)
{
    if (!box || !us_per_launch || pass < 0 || pass > 3 || reps < 1 || reps > 100000)
        return fail(SPH_ERR_ARG, "invalid argument (pass 0..3, reps 1..100000)");
    if (int rc = settle(box)) return rc;
    Arena &A = box->ar;
    HIPCHK(hipSetDevice(A.ctx->device));
    if (!A.stepped) return fail(SPH_ERR_ARG, "box has not been stepped (step it once before timing a pass)");
    // between sph_step calls the box rests in BOX_DONE, which the stepping
    // masks skip: the passes run under MASK_ALL, as the single evaluations
    // do.  They write only derived fields (never r, v, u, m), which the next
    // sph_step re-derives at its first evaluation.  The build stage (pass 3)
    // re-sorts the slots into cell order and rebuilds the lists at the
    // current positions: the particles keep their values, the next run builds
    // at its first step.
    const ArenaView V = A.view();
    const int f = A.fp32() ? 1 : 0;
    hipStream_t st = A.ctx->stream;
    A.host_dyn_ok = false;  // the passes may flag box health
    if (pass == 3) {
        A.sched_kept = false;
        if (int rc = A.build_now(MASK_ALL)) return rc;  // held builds settled before timing
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) e = hipEventRecord(ev[0], st);
    if (e == hipSuccess) {
        for (int i = 0; i < reps; ++i) {
            if (pass == 0) launch_density(V, f, MASK_ALL, A.unroll, st);
            else if (pass == 1) launch_grad(V, f, MASK_ALL, 0, A.unroll, st);
            else if (pass == 2) launch_rates(V, f, MASK_ALL, A.rates_unroll(), 0, st);
            else A.enqueue_build(MASK_ALL);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev[1], st);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return fail(SPH_ERR_HIP, std::string("sph_box_time_pass: ") + hipGetErrorString(e));
    *us_per_launch = 1000.0 * (double)ms / reps;
    return check_health(A, box->b);
}

// ---- batching and serving ------------------------------------------------------

int sph_arena_timing(sph_arena *ar, double *out, int n, int reset)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    return read_timing(ar->ar, out, n, reset);
}

int sph_arena_create(sph_ctx *ctx, int64_t particle_capacity, int precision, sph_arena **out)
{
    if (!ctx || !out || particle_capacity <= 0 || particle_capacity > (1ll << 30))
        return fail(SPH_ERR_ARG, "invalid argument");
    *out = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    sph_arena *a = new sph_arena();
    a->ar.ctx = ctx;
    // boxes of number density up to 2.5 (the vdW pole is at 1/b = 2) with
    // the default list radius 4.5
    int rc = a->ar.alloc((int)particle_capacity, precision, ARENA_LIST_ENTRIES);
    a->ar.always_build = 1;
    if (!rc) rc = a->ar.relayout();
    if (rc) {
        std::string keep = g_err;
        a->ar.release();
        delete a;
        g_err = keep;
        return rc;
    }
    *out = a;
    return SPH_OK;
}

int sph_arena_destroy(sph_arena *ar)
{
    if (!ar) return SPH_OK;
    (void)hipSetDevice(ar->ar.ctx->device);
    (void)hipStreamSynchronize(ar->ar.ctx->stream);
    ar->ar.release();
    delete ar;
    return SPH_OK;
}

int sph_arena_add(sph_arena *ar, int64_t n, const sph_params *prm, const double *r, const double *v,
                  const double *m, double T0, int64_t nsteps, int *box_id)
{
    if (!ar || !prm || !r || !v || !m || !box_id || nsteps < 0) return fail(SPH_ERR_ARG, "invalid argument");
    Arena &A = ar->ar;
    HIPCHK(hipSetDevice(A.ctx->device));
    for (int64_t i = 0; i < n; ++i) {
        const double x = r[2 * i], y = r[2 * i + 1];
        if (!(x >= 0.0 && x < prm->Lx && y >= 0.0 && y < prm->Ly))
            return fail(SPH_ERR_ARG, "position outside [0, L)");
    }
    int b = -1;
    if (int rc = A.add_box(n, prm, &b)) return rc;
    A.boxes[b].T0 = T0;
    HIPCHK(hipMemcpyAsync(A.d_boxes + b, &A.boxes[b], sizeof(BoxDesc), hipMemcpyHostToDevice, A.ctx->stream));
    if (int rc = A.upload(b, 0, r, 2)) return rc;
    if (int rc = A.upload(b, 1, v, 2)) return rc;
    if (int rc = A.upload(b, 2, m, 1)) return rc;
    BoxDyn D;
    memset(&D, 0, sizeof(D));
    D.state = BOX_INIT;
    D.nsteps = nsteps;
    if (int rc = A.write_dyn(b, D)) return rc;
    *box_id = b;
    return SPH_OK;
}

int sph_arena_step(sph_arena *ar, int64_t nsteps)
{
    if (!ar || nsteps < 0) return fail(SPH_ERR_ARG, "invalid argument");
    HIPCHK(hipSetDevice(ar->ar.ctx->device));
    return ar->ar.step(nsteps);
}

int sph_arena_poll(sph_arena *ar, int box_id, int64_t *steps_done, int *done)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    Arena &A = ar->ar;
    if (box_id < 0 || box_id >= (int)A.boxes.size() || !A.box_live[box_id]) return fail(SPH_ERR_ARG, "no such box");
    if (int rc = A.read_dyn()) return rc;
    const BoxDyn &D = A.h_dyn[box_id];
    if (steps_done) *steps_done = D.state == BOX_INIT ? -1 : (D.state == BOX_DONE ? D.nsteps : D.k);
    if (done) *done = D.state == BOX_DONE ? 1 : 0;
    if (D.health) return check_health(A, box_id);
    return SPH_OK;
}

int sph_arena_poll_all(sph_arena *ar, int nmax, int *state, int64_t *k, int *health)
{
    if (!ar || nmax < 0) return fail(SPH_ERR_ARG, "invalid argument");
    Arena &A = ar->ar;
    if (int rc = A.read_dyn()) return rc;
    const int nb = std::min(nmax, (int)A.boxes.size());
    for (int b = 0; b < nmax; ++b) {
        const bool live = b < nb && A.box_live[b];
        if (state) state[b] = live ? A.h_dyn[b].state : BOX_FREE;
        if (k) k[b] = live ? A.h_dyn[b].k : 0;
        if (health) health[b] = live ? A.h_dyn[b].health : 0;
    }
    return SPH_OK;
}

int sph_arena_result(sph_arena *ar, int box_id, double *r, double *v, double *u, double *rho, double *e0,
                     double *e1)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    Arena &A = ar->ar;
    if (box_id < 0 || box_id >= (int)A.boxes.size() || !A.box_live[box_id]) return fail(SPH_ERR_ARG, "no such box");
    HIPCHK(hipSetDevice(A.ctx->device));
    // r, v, u, rho packed into the device staging buffer, one copy to pinned
    // host memory and one wait (with the box states) for the four fields
    const int n = A.boxes[box_id].n;
    const size_t nd = (size_t)n;
    if (A.stage(sizeof(double) * 6 * nd)) return SPH_ERR_NOMEM;
    if (A.host_stage(sizeof(double) * 6 * nd)) return SPH_ERR_NOMEM;
    double *d = A.d_stage;
    const ArenaView V = A.view();
    launch_download(V, box_id, n, 0, d, A.ctx->stream);
    launch_download(V, box_id, n, 1, d + 2 * nd, A.ctx->stream);
    launch_download(V, box_id, n, 3, d + 4 * nd, A.ctx->stream);
    launch_download(V, box_id, n, 4, d + 5 * nd, A.ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(A.h_stage, d, sizeof(double) * 6 * nd, hipMemcpyDeviceToHost, A.ctx->stream));
    if (int rc = A.read_dyn()) return rc;  // waits for the copy above too
    const double *h = A.h_stage;
    if (r) memcpy(r, h, sizeof(double) * 2 * nd);
    if (v) memcpy(v, h + 2 * nd, sizeof(double) * 2 * nd);
    if (u) memcpy(u, h + 4 * nd, sizeof(double) * nd);
    if (rho) memcpy(rho, h + 5 * nd, sizeof(double) * nd);
    if (e0) *e0 = A.h_dyn[box_id].E0;
    if (e1) *e1 = A.h_dyn[box_id].E1;
    return SPH_OK;
}

int sph_arena_remove(sph_arena *ar, int box_id)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    HIPCHK(hipSetDevice(ar->ar.ctx->device));
    return ar->ar.remove_box(box_id);
}

int sph_arena_stats(sph_arena *ar, int64_t *used, int64_t *capacity, int *running, int *boxes)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    Arena &A = ar->ar;
    int64_t u = 0;
    int nb = 0;
    for (size_t b = 0; b < A.boxes.size(); ++b)
        if (A.box_live[b]) {
            u += A.boxes[b].n;
            ++nb;
        }
    if (used) *used = u;
    if (capacity) *capacity = A.cap;
    if (boxes) *boxes = nb;
    if (running) {
        if (int rc = A.read_dyn()) return rc;
        int nr = 0;
        for (size_t b = 0; b < A.boxes.size(); ++b)
            if (A.box_live[b] && (A.h_dyn[b].state == BOX_RUN || A.h_dyn[b].state == BOX_INIT)) ++nr;
        *running = nr;
    }
    return SPH_OK;
}

int sph_arena_set_graph(sph_arena *ar, int on)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    ar->ar.use_graph = on ? 1 : 0;
    ar->ar.invalidate_graph();
    return SPH_OK;
}

int sph_box_set_graph(sph_box *box, int on)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    box->ar.use_graph = on ? 1 : 0;
    box->ar.invalidate_graph();
    return box->ar.prepare_graphs();
}

int sph_host_alloc(int64_t bytes, void **out)
{
    if (!out || bytes < 0) return fail(SPH_ERR_ARG, "invalid argument");
    *out = nullptr;
    if (bytes == 0) return SPH_OK;
    if (hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(SPH_ERR_NOMEM, "hipHostMalloc failed");
    }
    return SPH_OK;
}

int sph_host_free(void *p)
{
    if (p) HIPCHK(hipHostFree(p));
    return SPH_OK;
}

int sph_box_snapshot_wait(sph_box *box)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    Arena &A = box->ar;
    if (A.snap_pending) {
        HIPCHK(hipEventSynchronize(A.ev_snap_done));
        A.snap_pending = false;
    }
    return SPH_OK;
}

int sph_box_snapshot(sph_box *box, double *r, double *v, double *u, double *m)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    if (int rc = settle(box)) return rc;
    Arena &A = box->ar;
    HIPCHK(hipSetDevice(A.ctx->device));
    if (int rc = sph_box_snapshot_wait(box)) return rc;  // the staging buffer is reused
    const int b = box->b, n = A.boxes[b].n;
    if (!A.copy_stream) {
        HIPCHK(hipStreamCreateWithFlags(&A.copy_stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&A.ev_snap_ready, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&A.ev_snap_done, hipEventDisableTiming));
    }
    const size_t need = (size_t)6 * n;
    if (need > A.snap_cap) {
        dfree(A.d_snap);
        A.snap_cap = 0;
        if (dalloc(&A.d_snap, need)) return SPH_ERR_NOMEM;
        A.snap_cap = need;
    }
    struct F { double *dst; int field, nc; size_t off; } fs[] = {
        {r, 0, 2, 0}, {v, 1, 2, (size_t)2 * n}, {u, 3, 1, (size_t)4 * n}, {m, 2, 1, (size_t)5 * n}};
    ArenaView V = A.view();
    for (auto &f : fs)
        if (f.dst) launch_download(V, b, n, f.field, A.d_snap + f.off, A.ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(A.ev_snap_ready, A.ctx->stream));
    HIPCHK(hipStreamWaitEvent(A.copy_stream, A.ev_snap_ready, 0));
    for (auto &f : fs)
        if (f.dst)
            HIPCHK(hipMemcpyAsync(f.dst, A.d_snap + f.off, sizeof(double) * f.nc * (size_t)n,
                                  hipMemcpyDeviceToHost, A.copy_stream));
    HIPCHK(hipEventRecord(A.ev_snap_done, A.copy_stream));
    A.snap_pending = true;
    return SPH_OK;
}

// Cell grid of the rendering: cells at least one cutoff wide, periods
// (xmax, ymax) as interpolate.py:79 uses them.
static int splash_grid(const double *bounds, double cutoff, SplashGrid &G)
{
    const double Lx = bounds[1], Ly = bounds[3];
    if (!(Lx > 0) || !(Ly > 0)) return fail(SPH_ERR_ARG, "bounds need xmax, ymax > 0");
    G.Lx = Lx; G.Ly = Ly; G.cutoff = cutoff;
    G.ncx = (int)std::floor(Lx / (cutoff * (1.0 + 1e-9)));
    G.ncy = (int)std::floor(Ly / (cutoff * (1.0 + 1e-9)));
    G.ncx = G.ncx < 3 ? 1 : std::min(G.ncx, 4096);
    G.ncy = G.ncy < 3 ? 1 : std::min(G.ncy, 4096);
    G.cwx = Lx / G.ncx; G.cwy = Ly / G.ncy;
    return SPH_OK;
}

int sph_splash_map_2d(sph_ctx *ctx, int64_t n, const double *r, const double *m, const double *rho,
                      const double *a, const double *h, int nx, const double *x, int ny, const double *y,
                      const double *bounds, double cutoff, double *Z)
{
    if (!ctx || n < 0 || nx < 1 || ny < 1 || !bounds || !x || !y || !Z || !(cutoff > 0) ||
        (n > 0 && (!r || !m || !rho || !a || !h)))
        return fail(SPH_ERR_ARG, "invalid argument");
    if (n > (1ll << 30) || (long long)nx * ny > (1ll << 31))
        return fail(SPH_ERR_ARG, "too many particles or pixels");
    SplashGrid G;
    if (int rc = splash_grid(bounds, cutoff, G)) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int ncell = G.ncx * G.ncy;
    const size_t np = (size_t)nx * ny;
    double *d_in = nullptr, *d_xy = nullptr, *d_Z = nullptr;
    double4 *d_rec = nullptr, *d_rec2 = nullptr;
    unsigned *d_keys = nullptr;
    int *d_cs = nullptr;
    void *d_tmp = nullptr;
    const size_t N = (size_t)std::max<int64_t>(n, 1);
    const size_t tmp_bytes = radix_sort_temp_bytes((int)N);
    int rc = 0;
    rc |= dalloc(&d_in, 6 * N);  // r (2n), m, rho, a, h
    rc |= dalloc(&d_xy, (size_t)nx + ny);
    rc |= dalloc(&d_Z, np);
    rc |= dalloc(&d_rec, N);
    rc |= dalloc(&d_rec2, N);
    rc |= dalloc(&d_keys, 5 * N);
    rc |= dalloc(&d_cs, (size_t)ncell + 1);
    rc |= dalloc((char **)&d_tmp, tmp_bytes);
    if (!rc) {
        const size_t b = sizeof(double) * n;
        if (n > 0) {
            if (hipMemcpyAsync(d_in, r, 2 * b, hipMemcpyHostToDevice, st) ||
                hipMemcpyAsync(d_in + 2 * n, m, b, hipMemcpyHostToDevice, st) ||
                hipMemcpyAsync(d_in + 3 * n, rho, b, hipMemcpyHostToDevice, st) ||
                hipMemcpyAsync(d_in + 4 * n, a, b, hipMemcpyHostToDevice, st) ||
                hipMemcpyAsync(d_in + 5 * n, h, b, hipMemcpyHostToDevice, st))
                rc = fail(SPH_ERR_HIP, "hipMemcpyAsync failed");
        }
        if (!rc && (hipMemcpyAsync(d_xy, x, sizeof(double) * nx, hipMemcpyHostToDevice, st) ||
                    hipMemcpyAsync(d_xy + nx, y, sizeof(double) * ny, hipMemcpyHostToDevice, st)))
            rc = fail(SPH_ERR_HIP, "hipMemcpyAsync failed");
        if (!rc) {
            unsigned *k_in = d_keys, *k_out = d_keys + N, *v_out = d_keys + 2 * N, *k_alt = d_keys + 3 * N,
                     *v_alt = d_keys + 4 * N;
            launch_splash(n, d_in, d_in + 2 * n, d_in + 3 * n, d_in + 4 * n, d_in + 5 * n, G, d_rec, d_rec2, k_in,
                          k_out, v_out, k_alt, v_alt, d_tmp, bits_for(ncell), d_cs, nx, ny, d_xy, d_xy + nx, d_Z,
                          st);
            if (hipGetLastError() != hipSuccess) rc = fail(SPH_ERR_HIP, "splash launch failed");
            if (!rc && hipMemcpyAsync(Z, d_Z, sizeof(double) * np, hipMemcpyDeviceToHost, st))
                rc = fail(SPH_ERR_HIP, "hipMemcpyAsync failed");
            if (!rc && hipStreamSynchronize(st)) rc = fail(SPH_ERR_HIP, "hipStreamSynchronize failed");
        }
    } else {
        rc = fail(SPH_ERR_NOMEM, "device allocation failed");
    }
    (void)hipStreamSynchronize(st);
    dfree(d_in); dfree(d_xy); dfree(d_Z); dfree(d_rec); dfree(d_rec2); dfree(d_keys); dfree(d_cs);
    if (d_tmp) (void)hipFree(d_tmp);
    return rc;
}

int sph_splash_map_2d_noalloc(sph_ctx *ctx, int64_t n, const double *r, const double *m, const double *rho,
                              const double *a, const double *h, int nx, const double *x, int ny, const double *y,
                              const double *bounds, double cutoff, double *Z)
{
    if (!ctx || n < 0 || nx < 1 || ny < 1 || !bounds || !x || !y || !Z || !(cutoff > 0) ||
        (n > 0 && (!r || !m || !rho || !a || !h)))
        return fail(SPH_ERR_ARG, "invalid argument");
    if (n > NOALLOC_MAX_PARTICLES)
        return fail(SPH_ERR_ARG, "sph_splash_map_2d_noalloc renders at most " + std::to_string(NOALLOC_MAX_PARTICLES) +
                                     " particles (its fixed scratch); use sph_splash_map_2d");
    if ((long long)nx * ny > (1ll << 31)) return fail(SPH_ERR_ARG, "too many pixels");
    SplashGrid G;
    if (int rc = splash_grid(bounds, cutoff, G)) return rc;
    const int ncell = G.ncx * G.ncy;
    if (ncell > NOALLOC_MAX_CELLS)
        return fail(SPH_ERR_ARG, "sph_splash_map_2d_noalloc holds at most " + std::to_string(NOALLOC_MAX_CELLS) +
                                     " grid cells (cells one cutoff wide); use sph_splash_map_2d");
    const int N = (int)std::max<int64_t>(n, 1);
    if (radix_sort_temp_bytes(N) > NOALLOC_SORT_BYTES) return fail(SPH_ERR_ARG, "sort counters exceed the scratch");
    unsigned char *base = (unsigned char *)g_noalloc_scratch[ctx->device];
    if (!base) return fail(SPH_ERR_STATE, "scratch address not set up for this device");
    HIPCHK(hipSetDevice(ctx->device));
    double4 *d_rec = (double4 *)base;
    double4 *d_rec2 = (double4 *)(base + NOALLOC_REC_BYTES);
    unsigned *d_keys = (unsigned *)(base + 2 * NOALLOC_REC_BYTES);
    int *d_cs = (int *)(base + 2 * NOALLOC_REC_BYTES + NOALLOC_KEY_BYTES);
    void *d_tmp = base + 2 * NOALLOC_REC_BYTES + NOALLOC_KEY_BYTES + NOALLOC_CS_BYTES;
    // every scratch word a launch reads it first writes (records, keys, the
    // cell starts of all ncell + 1 cells, the counters of every tile), so
    // nothing needs clearing between calls
    launch_splash(n, r, m, rho, a, h, G, d_rec, d_rec2, d_keys, d_keys + N, d_keys + 2 * N, d_keys + 3 * N,
                  d_keys + 4 * N, d_tmp, bits_for(ncell), d_cs, nx, ny, x, y, Z, ctx->stream);
    HIPCHK(hipGetLastError());
    return SPH_OK;
}

static int set_schedule(Arena &A, int unroll)
{
    if (unroll != 0 && unroll != 1 && unroll != 2 && unroll != 4)
        return fail(SPH_ERR_ARG, "unroll must be 0 (default), 1, 2 or 4");
    A.unroll = unroll;
    A.check_graph_sig();
    return SPH_OK;
}

int sph_arena_set_schedule(sph_arena *ar, int unroll)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    return set_schedule(ar->ar, unroll);
}

int sph_box_set_schedule(sph_box *box, int unroll)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    if (int rc = set_schedule(box->ar, unroll)) return rc;
    return box->ar.prepare_graphs();
}

static int set_rates_unroll(Arena &A, int unroll)
{
    if (unroll != 0 && unroll != 1 && unroll != 2 && unroll != 4)
        return fail(SPH_ERR_ARG, "rate-pass unroll must be 0 (as the other passes), 1, 2 or 4");
    A.unroll_rates = unroll;
    A.check_graph_sig();
    return SPH_OK;
}

static int set_pass_block(Arena &A, int block)
{
    if (block != 0 && block != 128 && block != 256 && block != 512)
        return fail(SPH_ERR_ARG, "pass block size must be 0 (default 256), 128, 256 or 512");
    A.pass_block = block ? block : 256;
    A.check_graph_sig();
    return SPH_OK;
}

int sph_arena_set_pass_block(sph_arena *ar, int block)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    return set_pass_block(ar->ar, block);
}

int sph_box_set_pass_block(sph_box *box, int block)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    if (int rc = set_pass_block(box->ar, block)) return rc;
    return box->ar.prepare_graphs();
}

// Entry width of the lists that the boxes' next builds write: 16 (slot
// offsets; a box whose offsets do not fit falls back to 32 by itself) or 32.
// The lists in memory keep their form until their box's next build.
static int set_list_bits(Arena &A, int bits)
{
    if (bits != 16 && bits != 32) return fail(SPH_ERR_ARG, "list entry bits must be 16 or 32");
    HIPCHK(hipSetDevice(A.ctx->device));
    A.list_bits = bits;
    for (size_t b = 0; b < A.boxes.size(); ++b)
        if (A.box_live[b]) A.boxes[b].build16 = bits == 16 ? 1 : 0;
    if (!A.boxes.empty())
        HIPCHK(hipMemcpyAsync(A.d_boxes, A.boxes.data(), sizeof(BoxDesc) * A.boxes.size(), hipMemcpyHostToDevice,
                              A.ctx->stream));
    return SPH_OK;
}

static int list_bits_of(Arena &A, int b, int *bits)
{
    if (!bits) return fail(SPH_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(A.ctx->device));
    if (int rc = A.read_dyn()) return rc;
    *bits = A.h_dyn[b].list16 ? 16 : 32;
    return SPH_OK;
}

int sph_arena_set_list_bits(sph_arena *ar, int bits)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    return set_list_bits(ar->ar, bits);
}

int sph_box_set_list_bits(sph_box *box, int bits)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    return set_list_bits(box->ar, bits);
}

int sph_arena_list_bits(sph_arena *ar, int box_id, int *bits)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    Arena &A = ar->ar;
    if (box_id < 0 || box_id >= (int)A.boxes.size() || !A.box_live[box_id]) return fail(SPH_ERR_ARG, "no such box");
    return list_bits_of(A, box_id, bits);
}

int sph_box_list_bits(sph_box *box, int *bits)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    return list_bits_of(box->ar, box->b, bits);
}

// Where the leapfrog step runs: inside the rate pass (1, the default) or as
// a launch of its own after it (0); the same results.
static int set_fused_step(Arena &A, int on)
{
    if (on != 0 && on != 1) return fail(SPH_ERR_ARG, "fused step must be 0 or 1");
    A.fuse_step = on;
    A.check_graph_sig();
    return SPH_OK;
}

int sph_arena_set_fused_step(sph_arena *ar, int on)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    return set_fused_step(ar->ar, on);
}

int sph_box_set_fused_step(sph_box *box, int on)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    if (int rc = set_fused_step(box->ar, on)) return rc;
    return box->ar.prepare_graphs();
}

int sph_arena_set_rates_unroll(sph_arena *ar, int unroll)
{
    if (!ar) return fail(SPH_ERR_ARG, "null arena");
    return set_rates_unroll(ar->ar, unroll);
}

int sph_box_set_rates_unroll(sph_box *box, int unroll)
{
    if (!box) return fail(SPH_ERR_ARG, "null box");
    if (int rc = set_rates_unroll(box->ar, unroll)) return rc;
    return box->ar.prepare_graphs();
}

int sph_arena_footprint(int64_t particle_capacity, int precision, int64_t *bytes)
{
    if (!bytes || particle_capacity < 0 || (precision != SPH_FP64 && precision != SPH_FP32C))
        return fail(SPH_ERR_ARG, "invalid argument");
    *bytes = (int64_t)arena_bytes((size_t)particle_capacity, precision, ARENA_LIST_ENTRIES);
    return SPH_OK;
}

}  // extern "C"
