// shll_api.cpp -- host side of the C ABI declared in include/shll.h.
//
// Owns device memory, one HIP stream per context or batch, HIP graphs of the step
// launches, the (instance, tile) work lists and the launch-shape autotuner.  No CPU
// compute path exists: every numerical operation on the state is a HIP kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "shll.h"
#include "shll_types.h"

namespace shll {
hipError_t launch_step1d(int prec, int order, int lim, int C, int wpb, int n_items, hipStream_t s, const StepArgs& a);
bool step1d_supported(int C, int wpb);
int step1d_items_per_block(int C, int wpb);
hipError_t launch_step2d(int prec, int order, int lim, int k, int wpb, int n_items, hipStream_t s,
                         const StepArgs& a);
bool step2d_supported(int prec, int order, int k);
hipError_t launch_init(int prec, int dim, int nblocks, int n_inst, hipStream_t st, const Inst* insts,
                       const double* prim, void* o0, void* o1, int64_t plane);
hipError_t launch_stage(int prec, int dim, int nblocks, hipStream_t st, const Inst* inst, const void* s0,
                        int64_t plane, double* out);
hipError_t launch_output(int prec, int dim, int nblocks, int n_inst, hipStream_t st, const Inst* insts,
                         const void* s0, const void* s1, int64_t plane, double* prim, double* cons);
hipError_t launch_sums(int prec, int dim, int nblocks, int n_inst, hipStream_t st, const Inst* insts,
                       const void* s0, const void* s1, int64_t plane, double* partial);
hipError_t launch_ctl(hipStream_t st, int64_t* ctl, int64_t v0, int64_t v1, int64_t v2, int add);
hipError_t launch_compare(int nblocks, hipStream_t st, const double* a, int64_t plane_a, const double* b,
                          int64_t plane_b, int64_t ncell, int dim, double* partial);
hipError_t noalloc_scratch(int device, double** out, size_t* capacity);
}  // namespace shll

using namespace shll;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static shll_status fail(shll_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

#define HIPCHECK(expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(SHLL_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                  \
    } while (0)

extern "C" const char* shll_last_error(void) { return g_err.c_str(); }
extern "C" const char* shll_version(void) { return "shll_amd 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// constants and step counts
// ---------------------------------------------------------------------------
static int ke_of(int dim, int order) { return (dim == 1 && order == 1) ? 0 : 1; }

// base_shll.c:17-22,83-84 (float) and the same in double; see Consts.
static Consts make_consts(int64_t nx, int64_t ny, double lx, double ly, double theta) {
    Consts k;
    std::memset(&k, 0, sizeof(k));
    {
        const float GAMMA = 1.4f, R = 1.0f, CFL = 0.25f;
        const float L = (float)lx, H = (float)ly;
        const float DX = L / (float)nx, DY = H / (float)ny;
        const float DT = (CFL / (R + 1.0f)) * DX;
        k.f_cv = (float)((double)R / ((double)GAMMA - 1.0));
        k.f_gR = GAMMA * R;
        k.f_R = R;
        k.f_lamx = CFL / (R + 1.0f);
        k.f_lamy = DT / DY;
    }
    {
        const double GAMMA = 1.4, R = 1.0, CFL = 0.25;
        const double DX = lx / (double)nx, DY = ly / (double)ny;
        const double DT = (CFL / (R + 1.0)) * DX;
        k.d_cv = R / (GAMMA - 1.0);
        k.d_gR = GAMMA * R;
        k.d_R = R;
        k.d_lamx = CFL / (R + 1.0);
        k.d_lamy = DT / DY;
        k.d_rcv = 1.0 / k.d_cv;
    }
    k.c_rcv = (float)(1.0 / k.d_cv);
    k.c_gR = (float)k.d_gR;
    k.c_R = (float)k.d_R;
    k.c_lamx = (float)k.d_lamx;
    k.c_lamy = (float)k.d_lamy;
    // MC weight: float in the reference (base-omp/2nd_order_base_shll.c:44)
    k.d_theta = theta;
    k.f_theta = (float)theta;
    k.c_theta = (float)theta;
    return k;
}

extern "C" int64_t shll_steps_to_time(const shll_problem* pb, float t_end) {
    if (!pb || pb->nx < 1) return -1;
    const float R = 1.0f, CFL = 0.25f;
    const float DX = (float)pb->lx / (float)pb->nx;
    const float DT = (CFL / (R + 1.0f)) * DX;
    float time = 0.0f;
    int64_t n = 0;
    while (time < t_end) {  // base_shll.c:208-218
        const float next = time + DT;
        // once DT is below half an ulp of time the float sum stops growing and the
        // reference's loop would never end: report that instead of spinning
        if (!(next > time)) return -1;
        time = next;
        n++;
    }
    return n;
}

static shll_status check_problem(const shll_problem* pb) {
    if (!pb) return fail(SHLL_EINVAL, "problem is NULL");
    if (pb->dim != 1 && pb->dim != 2) return fail(SHLL_EINVAL, "dim must be 1 or 2 (got %d)", pb->dim);
    if (pb->order != 1 && pb->order != 2) return fail(SHLL_EINVAL, "order must be 1 or 2 (got %d)", pb->order);
    if (pb->bc != SHLL_BC_REFLECT && pb->bc != SHLL_BC_OUTFLOW) return fail(SHLL_EINVAL, "bad bc %d", pb->bc);
    if (pb->precision < SHLL_F64 || pb->precision > SHLL_F64F)
        return fail(SHLL_EINVAL, "bad precision %d", pb->precision);
    if (pb->precision == SHLL_F64F && pb->dim != 1) return fail(SHLL_EINVAL, "SHLL_F64F is a 1D form");
    if (pb->nx < 2 || pb->nx > (1 << 30)) return fail(SHLL_EINVAL, "nx out of range (%lld)", (long long)pb->nx);
    if (pb->dim == 1 && pb->ny != 1) return fail(SHLL_EINVAL, "ny must be 1 in 1D");
    if (pb->dim == 2 && (pb->ny < 2 || pb->ny > (1 << 30))) return fail(SHLL_EINVAL, "ny out of range");
    if (pb->dim == 2 && (double)pb->nx * (double)pb->ny > 4.0e9) return fail(SHLL_EINVAL, "grid too large");
    if (pb->limiter != SHLL_LIM_MINMOD && pb->limiter != SHLL_LIM_MC)
        return fail(SHLL_EINVAL, "bad limiter %d", pb->limiter);
    if (pb->limiter == SHLL_LIM_MC && pb->order != 2)
        return fail(SHLL_EINVAL, "the MC limiter needs order 2");
    if (pb->limiter == SHLL_LIM_MC && !(pb->theta >= 1.0 && pb->theta <= 2.0))
        return fail(SHLL_EINVAL, "MC limiter theta must lie in [1, 2] (got %g)", pb->theta);
    if (!(pb->lx > 0.0) || !std::isfinite(pb->lx)) return fail(SHLL_EINVAL, "lx must be > 0");
    if (pb->dim == 2 && (!(pb->ly > 0.0) || !std::isfinite(pb->ly))) return fail(SHLL_EINVAL, "ly must be > 0");
    return SHLL_OK;
}

static shll_status ensure_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(SHLL_EDEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    return SHLL_OK;
}

static size_t word_bytes(int prec) { return prec == SHLL_F64 || prec == SHLL_F64F ? 8 : 4; }
static int nwords(int prec) { return prec == SHLL_C32 || prec == SHLL_C32K ? 2 : 1; }

// ---------------------------------------------------------------------------
// launch shapes and work lists
// ---------------------------------------------------------------------------
struct Launch {
    int C = 2;          // cells per lane (1D)
    int k = 8;          // steps per launch
    int wpb = 4;        // waves per block
    int graph = 128;    // steps per graph (0: no graph)
    int rows = 16;      // owned rows per 2D item
};

// Default owned rows per 2D item for k steps per launch: an item recomputes order*k halo
// rows on each side, so longer items waste less; 16 rows at one step per launch.  In 2D the
// launch's cells_per_lane sets the rows instead (each lane owns that many cells of its
// column); 0 or 1 selects this default.
static int rows2d(int k) { return k <= 1 ? 16 : (k == 2 ? 32 : 64); }
static constexpr int kMaxRows2d = 4096;
static constexpr int kMaxGraphLaunches = 4096;  // launches per captured graph

// 1D tiling of one instance of nx cells: segments of XW*64*C cells (XW wavefronts of a
// cooperative block, else 1), halo order*k.  The segment that holds a wall also holds
// that wall's ghost cell (-1 or nx), so the first segment starts at cell -1; the last
// one always reaches past cell nx.  A ghost cell takes its value from its neighbour by a
// lane shift within one wavefront, so ghost cell nx must not be a wavefront's first
// cell: such a segment starts one cell earlier.
static bool tile1d(int32_t inst, int64_t nx, int order, int C, int XW, int k, std::vector<Item1>& out) {
    const int64_t S = 64LL * C * XW;
    auto place_end = [&](int64_t g0) { return (nx - g0) % (64LL * C) == 0 ? g0 - 1 : g0; };
    if (nx + 2 <= S) {  // a shift by one then still fits: nx + 1 is a multiple of 64*C below S
        out.push_back(Item1{inst, (int32_t)place_end(-1), 0, (int32_t)nx});
        return true;
    }
    const int64_t H = (int64_t)order * k;
    if (S - 1 - H <= 0 || S - 2 * H <= 0) return false;
    // owned ranges start on 32-cell boundaries (128-byte lines of every plane, whose
    // instances start 64-cell aligned), so no line is written by two segments
    auto lines = [](int64_t w) { return w >= 32 ? w / 32 * 32 : w; };
    const int64_t W = lines(S - 2 * H);
    int64_t x0 = 0;
    {
        const int64_t w = std::min<int64_t>(nx, lines(S - 1 - H));
        out.push_back(Item1{inst, -1, 0, (int32_t)w});
        x0 = w;
    }
    while (x0 < nx) {
        const int64_t w = std::min<int64_t>(nx - x0, W);
        // only a segment that reaches past cell nx holds its ghost; one cell more of left
        // halo is harmless
        const int64_t g0 = x0 - H + S - 1 >= nx ? place_end(x0 - H) : x0 - H;
        if (g0 + S - 1 < std::min<int64_t>(x0 + w - 1 + H, nx)) return false;
        out.push_back(Item1{inst, (int32_t)g0, (int32_t)x0, (int32_t)w});
        x0 += w;
    }
    return true;
}

// 2D tiling: strips of 64 columns (one wavefront) whose outer H = order*k columns on a
// cut side are recomputed halo, times runs of `rows` owned rows (H halo rows each side).
static void tile2d(int32_t inst, int64_t nx, int64_t ny, int order, int k, int rows, std::vector<Item2>& out) {
    struct Strip { int32_t j0, jx0, jw; };
    std::vector<Strip> strips;
    const int64_t H = (int64_t)order * k;
    {
        int64_t jw = std::min<int64_t>(ny, 64 - H);
        strips.push_back(Strip{0, 0, (int32_t)jw});
        int64_t jx0 = jw;
        while (jx0 < ny) {
            int64_t w = std::min<int64_t>(ny - jx0, 64 - 2 * H);
            strips.push_back(Strip{(int32_t)(jx0 - H), (int32_t)jx0, (int32_t)w});
            jx0 += w;
        }
    }
    for (int64_t ix0 = 0; ix0 < nx; ix0 += rows) {
        const int64_t iw = std::min<int64_t>(rows, nx - ix0);
        for (const Strip& s : strips)
            out.push_back(Item2{inst, s.j0, s.jx0, s.jw, (int32_t)(ix0 - H), (int32_t)ix0, (int32_t)iw,
                                (int32_t)(iw + 2 * H)});
    }
}

// ---------------------------------------------------------------------------
// the engine shared by contexts (one instance) and batches (many)
// ---------------------------------------------------------------------------
struct Engine {
    int dim = 1, order = 1, prec = 0, nv = 3, lim = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t plane = 0;                 // cells per variable plane
    void* buf[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    void* init_copy[2] = {nullptr, nullptr};  // batch reset / autotune restore
    std::vector<Inst> h_inst;
    Inst* d_inst = nullptr;
    std::vector<int64_t> inst_cells;   // cells per instance
    std::vector<int64_t> order_idx;    // instances sorted by nsteps descending
    Launch L;
    // work list for the current launch shape (instances in order_idx order)
    void* d_items = nullptr;
    size_t d_items_cap = 0;
    int64_t n_items = 0;
    std::vector<int64_t> inst_first_item;  // prefix: items of instance rank r start at inst_first_item[r]
    bool items_valid = false;
    // progress
    int64_t launches_done = 0;  // launches since the initial state (their parity names the buffer of
                                // the instances still active)
    int64_t steps_done = 0;     // batch: steps every unfinished instance has taken
    std::vector<int64_t> launch_starts;  // batch: first step of every launch so far (ascending)
    int cur = 0;                // buffer holding the state of instances active now
    // graph (single-instance constant-parameter graphs, by buffer parity)
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int g_launches = 0;
    int g_key[3] = {0, 0, 0};
    // batch graphs: constant nodes whose launch index comes from the device counter
    // d_ctl (set before a run of chunks, advanced by each chunk's last node) and whose
    // item count comes from d_active[launch index]; one graph per (start parity, grid
    // blocks rounded up to a power of two)
    struct BatchGraph {
        hipGraphExec_t exec = nullptr;
        int parity = 0;
        int64_t blocks = 0;
    };
    std::vector<BatchGraph> bgs;
    int64_t* d_ctl = nullptr;           // {j0, anchor, jmax}, see StepArgs
    int32_t* d_active = nullptr;        // item counts of the anchored launch grid
    size_t active_cap = 0;
    int64_t anchor = -1;                // first step of the grid d_active describes (-1: none)
    std::vector<int32_t> h_active;      // host copy of d_active (source of its upload)
    // diagnostics scratch
    double* d_prim = nullptr;
    double* d_cons = nullptr;
    double* d_partial = nullptr;
    size_t partial_cap = 0;
    // timing
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_avg_ms = 0, last_total_ms = 0;
    int64_t last_launches = 0;
    bool batch = false;
};

static void engine_drop_graphs(Engine* E) {
    for (int p = 0; p < 2; ++p)
        if (E->gexec[p]) {
            (void)hipGraphExecDestroy(E->gexec[p]);
            E->gexec[p] = nullptr;
        }
    for (auto& bgx : E->bgs)
        if (bgx.exec) (void)hipGraphExecDestroy(bgx.exec);
    E->bgs.clear();
}

static void engine_free(Engine* E) {
    if (!E) return;
    (void)hipSetDevice(E->device);
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    engine_drop_graphs(E);
    for (int b = 0; b < 2; ++b)
        for (int w = 0; w < 2; ++w)
            if (E->buf[b][w]) (void)hipFree(E->buf[b][w]);
    for (int w = 0; w < 2; ++w)
        if (E->init_copy[w]) (void)hipFree(E->init_copy[w]);
    if (E->d_inst) (void)hipFree(E->d_inst);
    if (E->d_items) (void)hipFree(E->d_items);
    if (E->d_ctl) (void)hipFree(E->d_ctl);
    if (E->d_active) (void)hipFree(E->d_active);
    if (E->d_prim) (void)hipFree(E->d_prim);
    if (E->d_cons) (void)hipFree(E->d_cons);
    if (E->d_partial) (void)hipFree(E->d_partial);
    if (E->ev0) (void)hipEventDestroy(E->ev0);
    if (E->ev1) (void)hipEventDestroy(E->ev1);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
}

static int radius(const Engine* E) { return E->order; }

// steps this launch shape advances per launch
static int launch_k(const Engine* E) { return E->L.k; }

static int64_t active_items(const Engine* E, int64_t step_base);
static int64_t batch_total_steps(const Engine* E) {
    return E->order_idx.empty() ? 0 : E->h_inst[E->order_idx[0]].nsteps;
}

static shll_status engine_build_items(Engine* E) {
    if (E->items_valid) return SHLL_OK;
    // the graphs hold pointers to the tables rebuilt below: let any launch of them finish and
    // drop them before a table is freed or reallocated
    HIPCHECK(hipStreamSynchronize(E->stream));
    engine_drop_graphs(E);
    std::vector<Item1> it1;
    std::vector<Item2> it2;
    E->inst_first_item.assign(E->order_idx.size() + 1, 0);
    for (size_t r = 0; r < E->order_idx.size(); ++r) {
        const int64_t i = E->order_idx[r];
        const Inst& in = E->h_inst[i];
        E->inst_first_item[r] = E->dim == 1 ? (int64_t)it1.size() : (int64_t)it2.size();
        if (E->dim == 1) {
            if (!tile1d((int32_t)i, in.nx, E->order, E->L.C, E->L.wpb / step1d_items_per_block(E->L.C, E->L.wpb),
                        E->L.k, it1))
                return fail(SHLL_EINVAL, "launch shape C=%d k=%d leaves no owned cells (order %d)", E->L.C, E->L.k,
                            E->order);
        } else {
            tile2d((int32_t)i, in.nx, in.ny, E->order, E->L.k, E->L.rows, it2);
        }
    }
    E->inst_first_item[E->order_idx.size()] = E->dim == 1 ? (int64_t)it1.size() : (int64_t)it2.size();
    const size_t bytes = E->dim == 1 ? it1.size() * sizeof(Item1) : it2.size() * sizeof(Item2);
    E->n_items = E->dim == 1 ? (int64_t)it1.size() : (int64_t)it2.size();
    if (E->n_items > (int64_t)std::numeric_limits<int32_t>::max() / 2)
        return fail(SHLL_EINVAL, "too many work items");
    if (bytes > E->d_items_cap) {
        if (E->d_items) HIPCHECK(hipFree(E->d_items));
        E->d_items = nullptr;
        HIPCHECK(hipMalloc(&E->d_items, bytes));
        E->d_items_cap = bytes;
    }
    HIPCHECK(hipMemcpyAsync(E->d_items, E->dim == 1 ? (const void*)it1.data() : (const void*)it2.data(), bytes,
                            hipMemcpyHostToDevice, E->stream));
    HIPCHECK(hipStreamSynchronize(E->stream));
    if (E->batch) {
        // item counts of a grid of k-step launches (at most one launch per k steps of the
        // longest instance, whatever step it starts at), read by the constant graph nodes
        const int64_t jcap = batch_total_steps(E) / launch_k(E) + 2;
        if (jcap > (int64_t)std::numeric_limits<int32_t>::max())
            return fail(SHLL_EINVAL, "too many launches for one batch run");
        const size_t abytes = (size_t)jcap * sizeof(int32_t);
        if (abytes > E->active_cap) {
            if (E->d_active) HIPCHECK(hipFree(E->d_active));
            E->d_active = nullptr;
            HIPCHECK(hipMalloc(&E->d_active, abytes));
            E->active_cap = abytes;
        }
        if (!E->d_ctl) HIPCHECK(hipMalloc(&E->d_ctl, 3 * sizeof(int64_t)));
        E->anchor = -1;
    }
    E->items_valid = true;
    return SHLL_OK;
}

// number of items whose instance still has steps to take at step step_base (instances
// sorted by nsteps)
static int64_t active_items(const Engine* E, int64_t base) {
    // instances in order_idx are sorted by nsteps descending: find first with nsteps <= base
    size_t lo = 0, hi = E->order_idx.size();
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (E->h_inst[E->order_idx[mid]].nsteps > base) lo = mid + 1;
        else hi = mid;
    }
    return E->inst_first_item[lo];
}

static StepArgs make_args(const Engine* E, int64_t step_base, int parity, int64_t n_items) {
    StepArgs a;
    a.inst = E->d_inst;
    a.items = E->d_items;
    a.n_items = (int32_t)n_items;
    a.ksteps = launch_k(E);
    a.step_base = step_base;
    a.plane = E->plane;
    a.in0 = E->buf[parity][0];
    a.in1 = E->buf[parity][1];
    a.out0 = E->buf[parity ^ 1][0];
    a.out1 = E->buf[parity ^ 1][1];
    a.jctl = nullptr;
    a.active = nullptr;
    a.jofs = 0;
    return a;
}

// work items per block of a launch
static int items_per_block(const Engine* E) {
    return E->dim == 1 ? step1d_items_per_block(E->L.C, E->L.wpb) : E->L.wpb;
}

static hipError_t launch_one(const Engine* E, const StepArgs& a, int64_t n_items) {
    if (E->dim == 1) return launch_step1d(E->prec, E->order, E->lim, E->L.C, E->L.wpb, (int)n_items, E->stream, a);
    return launch_step2d(E->prec, E->order, E->lim, E->L.k, E->L.wpb, (int)n_items, E->stream, a);
}

// Single instance: every launch has the same items; ksteps varies only on the last one.
// Graph of g launches (g even) captured per buffer parity and replayed.
static int single_graph_launches(const Engine* E, int k) {
    return E->L.graph > 0 ? std::max(2, (E->L.graph / k) & ~1) : 0;
}

// The graph of g launches starting from buffer parity p: captured and instantiated once per
// launch shape (set_launch captures both parities, so a timed run never includes a capture)
static shll_status single_graph(Engine* E, int p, int g) {
    const int key[3] = {E->L.C, E->L.k, g};
    if (E->gexec[p] && std::memcmp(key, E->g_key, sizeof(key)) != 0) engine_drop_graphs(E);
    if (E->gexec[p]) return SHLL_OK;
    const int64_t n_items = E->n_items;
    hipGraph_t graph;
    HIPCHECK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    for (int q = 0; q < g; ++q) {
        StepArgs a = make_args(E, 0, p ^ (q & 1), n_items);
        hipError_t e = launch_one(E, a, n_items);
        if (e != hipSuccess) {
            hipGraph_t junk;
            (void)hipStreamEndCapture(E->stream, &junk);
            return fail(SHLL_EDEVICE, "launch during capture: %s", hipGetErrorString(e));
        }
    }
    HIPCHECK(hipStreamEndCapture(E->stream, &graph));
    hipError_t e = hipGraphInstantiate(&E->gexec[p], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(SHLL_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    std::memcpy(E->g_key, key, sizeof(key));
    return SHLL_OK;
}

static shll_status single_steps(Engine* E, int64_t nsteps) {
    const int k = launch_k(E);
    int64_t left = nsteps;
    const int64_t n_items = E->n_items;
    int64_t launches = 0;
    const int g = single_graph_launches(E, k);
    if (E->timing) HIPCHECK(hipEventRecord(E->ev0, E->stream));
    while (left > 0) {
        if (g > 0 && left >= (int64_t)g * k) {
            const int p = E->cur;
            shll_status st = single_graph(E, p, g);
            if (st != SHLL_OK) return st;
            HIPCHECK(hipGraphLaunch(E->gexec[p], E->stream));
            left -= (int64_t)g * k;
            launches += g;
            continue;  // g even: parity unchanged
        }
        StepArgs a = make_args(E, 0, E->cur, n_items);
        a.ksteps = (int32_t)std::min<int64_t>(k, left);
        HIPCHECK(launch_one(E, a, n_items));
        E->cur ^= 1;
        left -= a.ksteps;
        launches++;
    }
    if (E->timing) HIPCHECK(hipEventRecord(E->ev1, E->stream));
    HIPCHECK(hipStreamSynchronize(E->stream));
    E->launches_done += launches;
    E->last_launches = launches;
    if (E->timing && launches > 0) {
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
        E->last_total_ms = ms;
        E->last_avg_ms = ms / (double)launches;
    }
    return SHLL_OK;
}

// Batch: a launch at step s covers steps [s, s + r) (r <= k) of every instance with
// nsteps > s.  Whole chunks of g k-step launches replay constant graphs whose launches
// take their step base and item count from the device counter and the item-count table of
// a grid of k-step launches anchored at some step; the rest of a call is launched directly
// with host-side parameters, the last launch with the steps that remain.
static shll_status batch_graph(Engine* E, int parity, int64_t blocks, int g, hipGraphExec_t* out) {
    for (const auto& bgx : E->bgs)
        if (bgx.parity == parity && bgx.blocks == blocks) {
            *out = bgx.exec;
            return SHLL_OK;
        }
    hipGraph_t graph;
    HIPCHECK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int q = 0; q < g && e == hipSuccess; ++q) {
        StepArgs a = make_args(E, 0, (parity + q) & 1, 0);
        a.jctl = E->d_ctl;
        a.active = E->d_active;
        a.jofs = q;
        e = launch_one(E, a, blocks * items_per_block(E));
    }
    if (e == hipSuccess) e = launch_ctl(E->stream, E->d_ctl, g, 0, 0, 1);
    hipError_t e2 = hipStreamEndCapture(E->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(SHLL_EDEVICE, "batch graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    Engine::BatchGraph bgx;
    bgx.parity = parity;
    bgx.blocks = blocks;
    e = hipGraphInstantiate(&bgx.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(SHLL_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    E->bgs.push_back(bgx);
    *out = bgx.exec;
    return SHLL_OK;
}

// anchor the graph launches' grid at step s: the item count of every k-step launch from s
// to the end of the longest instance, uploaded to d_active (stream-ordered after every
// launch that read the previous table)
static shll_status batch_set_anchor(Engine* E, int64_t s) {
    const int k = launch_k(E);
    const int64_t jmax = (batch_total_steps(E) - s + k - 1) / k;
    E->h_active.assign((size_t)std::max<int64_t>(1, jmax), 0);
    if ((size_t)E->h_active.size() * sizeof(int32_t) > E->active_cap)
        return fail(SHLL_EINVAL, "internal: launch grid exceeds its table");
    for (int64_t j = 0; j < jmax; ++j) E->h_active[(size_t)j] = (int32_t)active_items(E, s + j * k);
    HIPCHECK(hipMemcpyAsync(E->d_active, E->h_active.data(), E->h_active.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice, E->stream));
    E->anchor = s;
    return SHLL_OK;
}

static shll_status batch_launches(Engine* E, int64_t nsteps) {
    const int k = launch_k(E);
    int64_t s = E->steps_done;
    const int64_t send = std::min<int64_t>(batch_total_steps(E), s + nsteps);
    const int g = E->L.graph > 0 ? std::max(2, (E->L.graph / k) & ~1) : 0;
    const int ipb = items_per_block(E);
    int64_t launches = 0;
    bool ctl_set = false;
    if (E->timing) HIPCHECK(hipEventRecord(E->ev0, E->stream));
    while (s < send) {
        if (g > 0 && send - s >= (int64_t)g * k) {
            if (!ctl_set) {
                if (E->anchor < 0 || E->anchor > s || (s - E->anchor) % k != 0) {
                    shll_status st = batch_set_anchor(E, s);
                    if (st != SHLL_OK) return st;
                }
                const int64_t jmax = (batch_total_steps(E) - E->anchor + k - 1) / k;
                HIPCHECK(launch_ctl(E->stream, E->d_ctl, (s - E->anchor) / k, E->anchor, jmax, 0));
                ctl_set = true;
            }
            // the chunk's first launch has the most active items; round its grid up to a
            // power of two so few graphs cover the whole run
            const int64_t need = (active_items(E, s) + ipb - 1) / ipb;
            int64_t blocks = 1;
            while (blocks < need) blocks <<= 1;
            hipGraphExec_t exec = nullptr;
            shll_status st = batch_graph(E, (int)((E->launches_done + launches) & 1), blocks, g, &exec);
            if (st != SHLL_OK) return st;
            HIPCHECK(hipGraphLaunch(exec, E->stream));
            for (int q = 0; q < g; ++q) E->launch_starts.push_back(s + (int64_t)q * k);
            s += (int64_t)g * k;
            launches += g;
            continue;
        }
        const int64_t na = active_items(E, s);
        StepArgs a = make_args(E, s, (int)((E->launches_done + launches) & 1), na);
        a.ksteps = (int32_t)std::min<int64_t>(k, send - s);
        HIPCHECK(launch_one(E, a, na));
        E->launch_starts.push_back(s);
        s += a.ksteps;
        launches++;
    }
    if (E->timing) HIPCHECK(hipEventRecord(E->ev1, E->stream));
    HIPCHECK(hipStreamSynchronize(E->stream));
    E->steps_done = s;
    E->launches_done += launches;
    E->cur = (int)(E->launches_done & 1);
    E->last_launches = launches;
    if (E->timing && launches > 0) {
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
        E->last_total_ms = ms;
        E->last_avg_ms = ms / (double)launches;
    }
    return SHLL_OK;
}

// back to the initial state's bookkeeping (the caller restores the state itself)
static void engine_rewind(Engine* E) {
    E->cur = 0;
    E->launches_done = 0;
    E->steps_done = 0;
    E->launch_starts.clear();
}

// buffer holding instance i's current state (batch bookkeeping)
static int inst_parity(const Engine* E, int64_t i) {
    if (!E->batch) return E->cur;
    // the launches instance i took part in: those that started before its last step
    const auto& st = E->launch_starts;
    const int64_t took = std::lower_bound(st.begin(), st.end(), E->h_inst[i].nsteps) - st.begin();
    return (int)(took & 1);
}

static shll_status engine_alloc(Engine* E, int64_t plane) {
    E->plane = plane;
    const size_t bytes = (size_t)plane * E->nv * word_bytes(E->prec);
    for (int b = 0; b < 2; ++b)
        for (int w = 0; w < nwords(E->prec); ++w) {
            hipError_t e = hipMalloc(&E->buf[b][w], bytes);
            if (e != hipSuccess) return fail(SHLL_ENOMEM, "hipMalloc(%zu) for the state: %s", bytes, hipGetErrorString(e));
            HIPCHECK(hipMemsetAsync(E->buf[b][w], 0, bytes, E->stream));
        }
    const size_t inst_bytes = E->h_inst.size() * sizeof(Inst);
    HIPCHECK(hipMalloc(&E->d_inst, inst_bytes));
    HIPCHECK(hipMemcpyAsync(E->d_inst, E->h_inst.data(), inst_bytes, hipMemcpyHostToDevice, E->stream));
    HIPCHECK(hipEventCreate(&E->ev0));
    HIPCHECK(hipEventCreate(&E->ev1));
    return SHLL_OK;
}

static int elem_blocks(int64_t ncell) { return (int)std::max<int64_t>(1, std::min<int64_t>(2048, (ncell + 255) / 256)); }

// upload primitives of all instances and convert them (k_init) into buffer 0
static shll_status engine_init_state(Engine* E, const double* const* prims) {
    const size_t pbytes = (size_t)E->plane * E->nv * sizeof(double);
    double* d_p = nullptr;
    HIPCHECK(hipMalloc(&d_p, pbytes));
    // gather every instance's primitives into one pinned staging buffer in the device
    // layout, then one host-to-device copy (instead of nvar copies per instance)
    double* h_p = nullptr;
    if (hipHostMalloc((void**)&h_p, pbytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(d_p);
        return fail(SHLL_ENOMEM, "hipHostMalloc(%zu) for the initial data", pbytes);
    }
    int64_t maxcells = 0;
    for (size_t i = 0; i < E->h_inst.size(); ++i) {
        const int64_t nc = E->inst_cells[i];
        maxcells = std::max(maxcells, nc);
        for (int v = 0; v < E->nv; ++v)
            std::memcpy(h_p + v * E->plane + E->h_inst[i].off, prims[i] + (size_t)v * nc, (size_t)nc * sizeof(double));
    }
    hipError_t eu = hipMemcpyAsync(d_p, h_p, pbytes, hipMemcpyHostToDevice, E->stream);
    hipError_t es = hipStreamSynchronize(E->stream);
    (void)hipHostFree(h_p);
    if (eu != hipSuccess || es != hipSuccess) {
        (void)hipFree(d_p);
        return fail(SHLL_EDEVICE, "upload: %s", hipGetErrorString(eu != hipSuccess ? eu : es));
    }
    hipError_t e = launch_init(E->prec, E->dim, elem_blocks(maxcells), (int)E->h_inst.size(), E->stream, E->d_inst, d_p,
                               E->buf[0][0], E->buf[0][1], E->plane);
    hipError_t e2 = hipStreamSynchronize(E->stream);
    (void)hipFree(d_p);
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(SHLL_EDEVICE, "init kernel: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    engine_rewind(E);
    return SHLL_OK;
}

static shll_status engine_snapshot(Engine* E, bool save) {
    const size_t bytes = (size_t)E->plane * E->nv * word_bytes(E->prec);
    for (int w = 0; w < nwords(E->prec); ++w) {
        if (save) {
            if (!E->init_copy[w]) HIPCHECK(hipMalloc(&E->init_copy[w], bytes));
            HIPCHECK(hipMemcpyAsync(E->init_copy[w], E->buf[E->cur][w], bytes, hipMemcpyDeviceToDevice, E->stream));
        } else {
            if (!E->init_copy[w]) return fail(SHLL_ESTATE, "no snapshot to restore");
            HIPCHECK(hipMemcpyAsync(E->buf[0][w], E->init_copy[w], bytes, hipMemcpyDeviceToDevice, E->stream));
        }
    }
    HIPCHECK(hipStreamSynchronize(E->stream));
    return SHLL_OK;
}

// primitives (and conserved) of instance i into host arrays (nv * cells each)
static shll_status engine_get(Engine* E, int64_t i, double* prim, double* cons) {
    const int64_t nc = E->inst_cells[i];
    if (!E->d_prim) HIPCHECK(hipMalloc(&E->d_prim, (size_t)E->plane * E->nv * sizeof(double)));
    if (cons && !E->d_cons) HIPCHECK(hipMalloc(&E->d_cons, (size_t)E->plane * E->nv * sizeof(double)));
    const int p = inst_parity(E, i);
    // one instance: point the kernel at its Inst record
    HIPCHECK(launch_output(E->prec, E->dim, elem_blocks(nc), 1, E->stream, E->d_inst + i, E->buf[p][0], E->buf[p][1],
                           E->plane, E->d_prim, cons ? E->d_cons : nullptr));
    for (int v = 0; v < E->nv; ++v) {
        const size_t off = (size_t)v * E->plane + E->h_inst[i].off;
        if (prim) HIPCHECK(hipMemcpyAsync(prim + (size_t)v * nc, E->d_prim + off, nc * sizeof(double),
                                          hipMemcpyDeviceToHost, E->stream));
        if (cons) HIPCHECK(hipMemcpyAsync(cons + (size_t)v * nc, E->d_cons + off, nc * sizeof(double),
                                          hipMemcpyDeviceToHost, E->stream));
    }
    HIPCHECK(hipStreamSynchronize(E->stream));
    return SHLL_OK;
}

static shll_status ensure_partial(Engine* E, size_t n) {
    if (n <= E->partial_cap) return SHLL_OK;
    if (E->d_partial) HIPCHECK(hipFree(E->d_partial));
    E->d_partial = nullptr;
    HIPCHECK(hipMalloc(&E->d_partial, n * sizeof(double)));
    E->partial_cap = n;
    return SHLL_OK;
}

// FP64 sums per instance (out[i*5 + q]: q < nv sums, q = 4 bad-cell count)
static shll_status engine_sums(Engine* E, std::vector<double>& out) {
    const int64_t ninst = (int64_t)E->h_inst.size();
    int64_t maxcells = 0;
    for (int64_t c : E->inst_cells) maxcells = std::max(maxcells, c);
    const int nb = std::min(elem_blocks(maxcells), 256);
    shll_status st = ensure_partial(E, (size_t)ninst * nb * 6);
    if (st != SHLL_OK) return st;
    out.assign((size_t)ninst * 5, 0.0);
    std::vector<double> part((size_t)ninst * nb * 6);
    // instances may sit in different buffers: launch per parity group
    for (int p = 0; p < 2; ++p) {
        HIPCHECK(launch_sums(E->prec, E->dim, nb, (int)ninst, E->stream, E->d_inst, E->buf[p][0], E->buf[p][1],
                             E->plane, E->d_partial));
        HIPCHECK(hipMemcpyAsync(part.data(), E->d_partial, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                E->stream));
        HIPCHECK(hipStreamSynchronize(E->stream));
        for (int64_t i = 0; i < ninst; ++i) {
            if (inst_parity(E, i) != p) continue;
            for (int q = 0; q < 5; ++q) {
                double t = 0;
                for (int b = 0; b < nb; ++b) t += part[((size_t)i * nb + b) * 6 + q];
                out[(size_t)i * 5 + q] = t;
            }
        }
    }
    return SHLL_OK;
}

// ---------------------------------------------------------------------------
// single-instance context
// ---------------------------------------------------------------------------
struct shll_ctx {
    shll_problem pb;
    Engine* E;
};

static Engine* engine_new(int dim, int order, int prec, int lim) {
    Engine* E = new Engine();
    E->dim = dim;
    E->order = order;
    E->prec = prec;
    E->lim = lim;
    E->nv = dim == 1 ? 3 : 4;
    return E;
}

static shll_status engine_create_stream(Engine* E) {
    HIPCHECK(hipGetDevice(&E->device));
    HIPCHECK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    return SHLL_OK;
}

static void default_launch(Engine* E, int64_t max_cells) {
    if (E->dim == 2) {
        // First order: several steps per launch (row-pipelined stages); measured on a
        // 4096^2 domain and a 64-domain batch, the FP32 forms peak at 4 steps per launch,
        // FP64 (twice the registers per stage) at 2.  Second order: C32K 2 steps per launch
        // (7.5e10 vs 5.9e10 cell-updates/s at 4096^2); the TwoSum form's two-row stages
        // overflow the registers (4.4e10 vs 5.3e10), so it and FP64/REF32 take one step per
        // launch in 16-row items (a 4096^2 domain in 64-row items leaves ~4 wavefronts per
        // SIMD, too few to cover the row loads).
        E->L.C = 1;
        E->L.k = E->order == 1 ? (E->prec == PREC_C32 || E->prec == PREC_C32K ? 4 : 2) : (E->prec == PREC_C32K ? 2 : 1);
        E->L.wpb = 4;
        E->L.rows = rows2d(E->L.k);
        E->L.graph = 64;
        return;
    }
    // small instances: short segments and a few steps per launch keep many wavefronts busy
    E->L.C = max_cells <= 4096 ? 1 : 2;
    E->L.k = max_cells <= 4096 ? 8 : 16;
    E->L.wpb = max_cells <= 65536 ? 1 : 4;
    E->L.graph = 256;
}

extern "C" shll_status shll_init(const shll_problem* pb, const double* prim_soa, shll_ctx** out) {
    if (!out) return fail(SHLL_EINVAL, "out is NULL");
    *out = nullptr;
    shll_status st = check_problem(pb);
    if (st != SHLL_OK) return st;
    if (!prim_soa) return fail(SHLL_EINVAL, "prim_soa is NULL");
    if ((st = ensure_device()) != SHLL_OK) return st;
    Engine* E = engine_new(pb->dim, pb->order, pb->precision, pb->limiter);
    if ((st = engine_create_stream(E)) != SHLL_OK) { engine_free(E); return st; }
    Inst in;
    std::memset(&in, 0, sizeof(in));
    in.off = 0;
    in.nsteps = std::numeric_limits<int64_t>::max() / 4;
    in.nx = (int32_t)pb->nx;
    in.ny = (int32_t)pb->ny;
    in.bc = pb->bc;
    in.ke = ke_of(pb->dim, pb->order);
    in.k = make_consts(pb->nx, pb->ny, pb->lx, pb->dim == 2 ? pb->ly : 1.0, pb->theta);
    E->h_inst.push_back(in);
    const int64_t cells = pb->nx * pb->ny;
    E->inst_cells.push_back(cells);
    E->order_idx.push_back(0);
    default_launch(E, cells);
    if ((st = engine_alloc(E, (cells + 63) / 64 * 64)) != SHLL_OK) { engine_free(E); return st; }
    const double* prims[1] = {prim_soa};
    if ((st = engine_init_state(E, prims)) != SHLL_OK) { engine_free(E); return st; }
    if ((st = engine_build_items(E)) != SHLL_OK) { engine_free(E); return st; }
    shll_ctx* c = new shll_ctx;
    c->pb = *pb;
    c->E = E;
    *out = c;
    return SHLL_OK;
}

extern "C" void shll_free(shll_ctx* ctx) {
    if (!ctx) return;
    engine_free(ctx->E);
    delete ctx;
}

extern "C" shll_status shll_step(shll_ctx* ctx, int64_t nsteps) {
    if (!ctx) return fail(SHLL_EINVAL, "ctx is NULL");
    if (nsteps < 0) return fail(SHLL_EINVAL, "nsteps < 0");
    HIPCHECK(hipSetDevice(ctx->E->device));
    shll_status st = engine_build_items(ctx->E);
    if (st != SHLL_OK) return st;
    return single_steps(ctx->E, nsteps);
}

extern "C" shll_status shll_run_to_time(shll_ctx* ctx, float t_end, int64_t* steps_out) {
    if (!ctx) return fail(SHLL_EINVAL, "ctx is NULL");
    const int64_t n = shll_steps_to_time(&ctx->pb, t_end);
    if (n < 0)
        return fail(SHLL_EINVAL, "the float time loop never reaches t_end = %g: DT is below half an ulp of the "
                                 "time before it gets there", (double)t_end);
    shll_status st = shll_step(ctx, n);
    if (st != SHLL_OK) return st;
    if (steps_out) *steps_out = n;
    return shll_check_state(ctx, nullptr);
}

extern "C" shll_status shll_get_primitives(shll_ctx* ctx, double* out_soa) {
    if (!ctx || !out_soa) return fail(SHLL_EINVAL, "NULL argument");
    HIPCHECK(hipSetDevice(ctx->E->device));
    return engine_get(ctx->E, 0, out_soa, nullptr);
}

extern "C" shll_status shll_get_conserved(shll_ctx* ctx, double* out_soa) {
    if (!ctx || !out_soa) return fail(SHLL_EINVAL, "NULL argument");
    HIPCHECK(hipSetDevice(ctx->E->device));
    return engine_get(ctx->E, 0, nullptr, out_soa);
}

extern "C" size_t shll_state_bytes(const shll_ctx* ctx) {
    if (!ctx) return 0;
    const Engine* E = ctx->E;
    return (size_t)E->inst_cells[0] * E->nv * word_bytes(E->prec) * nwords(E->prec);
}

extern "C" shll_status shll_get_state_raw(shll_ctx* ctx, void* out, size_t bytes) {
    if (!ctx || !out) return fail(SHLL_EINVAL, "NULL argument");
    if (bytes != shll_state_bytes(ctx)) return fail(SHLL_EINVAL, "bytes must be %zu", shll_state_bytes(ctx));
    Engine* E = ctx->E;
    HIPCHECK(hipSetDevice(E->device));
    const size_t wb = word_bytes(E->prec);
    const int64_t nc = E->inst_cells[0];
    char* o = (char*)out;
    for (int w = 0; w < nwords(E->prec); ++w)
        for (int v = 0; v < E->nv; ++v) {
            HIPCHECK(hipMemcpyAsync(o, (const char*)E->buf[E->cur][w] + (size_t)v * E->plane * wb, nc * wb,
                                    hipMemcpyDeviceToHost, E->stream));
            o += nc * wb;
        }
    HIPCHECK(hipStreamSynchronize(E->stream));
    return SHLL_OK;
}

extern "C" shll_status shll_set_state_raw(shll_ctx* ctx, const void* in, size_t bytes) {
    if (!ctx || !in) return fail(SHLL_EINVAL, "NULL argument");
    if (bytes != shll_state_bytes(ctx)) return fail(SHLL_EINVAL, "bytes must be %zu", shll_state_bytes(ctx));
    Engine* E = ctx->E;
    HIPCHECK(hipSetDevice(E->device));
    const size_t wb = word_bytes(E->prec);
    const int64_t nc = E->inst_cells[0];
    const char* p = (const char*)in;
    for (int w = 0; w < nwords(E->prec); ++w)
        for (int v = 0; v < E->nv; ++v) {
            HIPCHECK(hipMemcpyAsync((char*)E->buf[E->cur][w] + (size_t)v * E->plane * wb, p, nc * wb,
                                    hipMemcpyHostToDevice, E->stream));
            p += nc * wb;
        }
    HIPCHECK(hipStreamSynchronize(E->stream));
    return SHLL_OK;
}

// Save_Results (base_shll.c:180-192; base_shll_2d.c:322-340): values pass through float
// as in the reference's arrays; cx = (i+0.5)*DX evaluated in double and stored to float.
extern "C" shll_status shll_write_results(shll_ctx* ctx, const char* path) {
    if (!ctx || !path) return fail(SHLL_EINVAL, "NULL argument");
    const shll_problem& pb = ctx->pb;
    const int64_t nc = pb.nx * pb.ny;
    const int nv = pb.dim == 1 ? 3 : 4;
    std::vector<double> p((size_t)nc * nv);
    shll_status st = shll_get_primitives(ctx, p.data());
    if (st != SHLL_OK) return st;
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(SHLL_EINVAL, "cannot open %s", path);
    const float DX = (float)pb.lx / (float)pb.nx;
    if (pb.dim == 1) {
        for (int64_t i = 0; i < pb.nx; ++i) {
            const float cx = (float)(((double)i + 0.5) * (double)DX);
            std::fprintf(f, "%e\t%e\t%e\t%e\n", (double)cx, (double)(float)p[i], (double)(float)p[nc + i],
                         (double)(float)p[2 * nc + i]);
        }
    } else {
        const float DY = (float)pb.ly / (float)pb.ny;
        for (int64_t i = 0; i < pb.nx; ++i)
            for (int64_t j = 0; j < pb.ny; ++j) {
                const int64_t c = i * pb.ny + j;
                const float cx = (float)(((double)i + 0.5) * (double)DX);
                const float cy = (float)(((double)j + 0.5) * (double)DY);
                std::fprintf(f, "%e\t%e\t%e\t%e\t%e\t%e\n", (double)cx, (double)cy, (double)(float)p[c],
                             (double)(float)p[nc + c], (double)(float)p[2 * nc + c], (double)(float)p[3 * nc + c]);
            }
    }
    if (std::fclose(f) != 0) return fail(SHLL_EINVAL, "write to %s failed", path);
    return SHLL_OK;
}

// The omp-vectors Save_Results (2nd_order_arm_sve.c:921-940): primitives and conserved
// variables of every cell, as floats, cell centres (i+0.5)*DX computed in double.
extern "C" shll_status shll_write_results_full(shll_ctx* ctx, const char* path) {
    if (!ctx || !path) return fail(SHLL_EINVAL, "NULL argument");
    const shll_problem& pb = ctx->pb;
    const int64_t nc = pb.nx * pb.ny;
    const int nv = pb.dim == 1 ? 3 : 4;
    std::vector<double> p((size_t)nc * nv), u((size_t)nc * nv);
    shll_status st = shll_get_primitives(ctx, p.data());
    if (st == SHLL_OK) st = shll_get_conserved(ctx, u.data());
    if (st != SHLL_OK) return st;
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(SHLL_EINVAL, "cannot open %s", path);
    const float DX = (float)pb.lx / (float)pb.nx;
    const float DY = pb.dim == 2 ? (float)pb.ly / (float)pb.ny : 1.0f;
    for (int64_t i = 0; i < pb.nx; ++i)
        for (int64_t j = 0; j < pb.ny; ++j) {
            const int64_t c = i * pb.ny + j;
            const float cx = (float)(((double)i + 0.5) * (double)DX);
            std::fprintf(f, "%e", (double)cx);
            if (pb.dim == 2) std::fprintf(f, "\t%e", (double)(float)(((double)j + 0.5) * (double)DY));
            for (int v = 0; v < nv; ++v) std::fprintf(f, "\t%e", (double)(float)p[(size_t)v * nc + c]);
            for (int v = 0; v < nv; ++v) std::fprintf(f, "\t%e", (double)(float)u[(size_t)v * nc + c]);
            std::fputc('\n', f);
        }
    if (std::fclose(f) != 0) return fail(SHLL_EINVAL, "write to %s failed", path);
    return SHLL_OK;
}

static double cell_volume(const shll_problem& pb) {
    return pb.dim == 1 ? pb.lx / (double)pb.nx : (pb.lx / (double)pb.nx) * (pb.ly / (double)pb.ny);
}

extern "C" shll_status shll_conserved_sums(shll_ctx* ctx, double* out) {
    if (!ctx || !out) return fail(SHLL_EINVAL, "NULL argument");
    HIPCHECK(hipSetDevice(ctx->E->device));
    std::vector<double> s;
    shll_status st = engine_sums(ctx->E, s);
    if (st != SHLL_OK) return st;
    const double vol = cell_volume(ctx->pb);
    for (int v = 0; v < ctx->E->nv; ++v) out[v] = s[v] * vol;
    return SHLL_OK;
}

// shll_conserved_sums without any allocation: the same kernel, grid and summation order,
// with the block partials in the library's static device scratch and on this stack
extern "C" shll_status shll_conserved_sums_noalloc(shll_ctx* ctx, double* out) {
    if (!ctx || !out) return fail(SHLL_EINVAL, "NULL argument");
    Engine* E = ctx->E;
    HIPCHECK(hipSetDevice(E->device));
    double* scratch = nullptr;
    size_t cap = 0;
    HIPCHECK(noalloc_scratch(E->device, &scratch, &cap));
    constexpr int kMaxBlocks = 256;
    const int nb = std::min(elem_blocks(E->inst_cells[0]), kMaxBlocks);
    if ((size_t)nb * 6 > cap)
        return fail(SHLL_EINVAL, "%d partial blocks exceed the %zu-double scratch of the allocation-free sums", nb,
                    cap);
    HIPCHECK(launch_sums(E->prec, E->dim, nb, 1, E->stream, E->d_inst, E->buf[E->cur][0], E->buf[E->cur][1],
                         E->plane, scratch));
    double part[kMaxBlocks * 6];
    HIPCHECK(hipMemcpyAsync(part, scratch, (size_t)nb * 6 * sizeof(double), hipMemcpyDeviceToHost, E->stream));
    HIPCHECK(hipStreamSynchronize(E->stream));
    const double vol = cell_volume(ctx->pb);
    for (int v = 0; v < E->nv; ++v) {
        double t = 0;
        for (int b = 0; b < nb; ++b) t += part[(size_t)b * 6 + v];
        out[v] = t * vol;
    }
    return SHLL_OK;
}

extern "C" shll_status shll_check_state(shll_ctx* ctx, int64_t* bad_cells) {
    if (!ctx) return fail(SHLL_EINVAL, "ctx is NULL");
    HIPCHECK(hipSetDevice(ctx->E->device));
    std::vector<double> s;
    shll_status st = engine_sums(ctx->E, s);
    if (st != SHLL_OK) return st;
    const int64_t bad = (int64_t)s[4];
    if (bad_cells) *bad_cells = bad;
    if (bad > 0) return fail(SHLL_ENONFINITE, "%lld cells with non-finite or non-positive density/energy", (long long)bad);
    return SHLL_OK;
}

extern "C" shll_status shll_stage_fluxes(shll_ctx* ctx, double* a, double* fp, double* fm, double* hp, double* hm) {
    if (!ctx) return fail(SHLL_EINVAL, "ctx is NULL");
    Engine* E = ctx->E;
    if (E->dim == 1 && (hp || hm)) return fail(SHLL_EINVAL, "y fluxes requested for a 1D problem");
    HIPCHECK(hipSetDevice(E->device));
    const int64_t nc = E->inst_cells[0], P = E->plane;
    double* d = nullptr;
    HIPCHECK(hipMalloc(&d, (size_t)(1 + 4 * E->nv) * P * sizeof(double)));
    hipError_t e = launch_stage(E->prec, E->dim, elem_blocks(nc), E->stream, E->d_inst, E->buf[E->cur][0], P, d);
    double* outs[5] = {a, fp, fm, hp, hm};
    for (int k = 0; k < 5 && e == hipSuccess; ++k) {
        if (!outs[k]) continue;
        const int nv = k == 0 ? 1 : E->nv;
        const double* src = k == 0 ? d : d + P + (size_t)(k - 1) * E->nv * P;
        for (int v = 0; v < nv && e == hipSuccess; ++v)
            e = hipMemcpyAsync(outs[k] + (size_t)v * nc, src + (size_t)v * P, nc * sizeof(double),
                               hipMemcpyDeviceToHost, E->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(E->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SHLL_EDEVICE, "stage evaluation: %s", hipGetErrorString(e));
    return SHLL_OK;
}

// primitives of instance i on the device (E->d_prim at the instance's offset)
static shll_status engine_prim_dev(Engine* E, int64_t i, const double** out) {
    if (!E->d_prim) HIPCHECK(hipMalloc(&E->d_prim, (size_t)E->plane * E->nv * sizeof(double)));
    const int p = inst_parity(E, i);
    HIPCHECK(launch_output(E->prec, E->dim, elem_blocks(E->inst_cells[i]), 1, E->stream, E->d_inst + i,
                           E->buf[p][0], E->buf[p][1], E->plane, E->d_prim, nullptr));
    HIPCHECK(hipStreamSynchronize(E->stream));
    *out = E->d_prim + E->h_inst[i].off;
    return SHLL_OK;
}

static shll_status compare_dev(Engine* E, int dim, const double* a, int64_t plane_a, const double* b,
                               int64_t plane_b, int64_t ncell, double vol, shll_errors* out) {
    const int nb = std::min(elem_blocks(ncell), 512);
    shll_status st = ensure_partial(E, (size_t)nb * 10);
    if (st != SHLL_OK) return st;
    HIPCHECK(launch_compare(nb, E->stream, a, plane_a, b, plane_b, ncell, dim, E->d_partial));
    std::vector<double> part((size_t)nb * 10);
    HIPCHECK(hipMemcpyAsync(part.data(), E->d_partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, E->stream));
    HIPCHECK(hipStreamSynchronize(E->stream));
    double s[4] = {0, 0, 0, 0}, mx[6] = {0, 0, 0, 0, 0, 0};
    for (int bb = 0; bb < nb; ++bb) {
        for (int q = 0; q < 4; ++q) s[q] += part[(size_t)bb * 10 + q];
        for (int q = 0; q < 6; ++q) mx[q] = std::max(mx[q], part[(size_t)bb * 10 + 4 + q]);
    }
    out->l1_rho = s[0] * vol;
    out->l1_u = s[1] * vol;
    out->l1_p = s[2] * vol;
    out->nonfinite = s[3];
    out->max_rel_rho = mx[0];
    out->max_rel_p = mx[1];
    out->max_du = mx[2];
    out->max_abs_u_ref = mx[3];
    out->linf_rho = mx[4];
    out->linf_p = mx[5];
    return SHLL_OK;
}

extern "C" shll_status shll_compare(shll_ctx* ctx, shll_ctx* ref, shll_errors* out) {
    if (!ctx || !ref || !out) return fail(SHLL_EINVAL, "NULL argument");
    if (ctx->pb.dim != ref->pb.dim || ctx->pb.nx != ref->pb.nx || ctx->pb.ny != ref->pb.ny)
        return fail(SHLL_EINVAL, "contexts have different shapes");
    if (ctx->E->device != ref->E->device) return fail(SHLL_EINVAL, "contexts on different devices");
    HIPCHECK(hipSetDevice(ctx->E->device));
    Engine *A = ctx->E, *B = ref->E;
    const double *pa = nullptr, *pb = nullptr;
    shll_status st = engine_prim_dev(A, 0, &pa);
    if (st == SHLL_OK) st = engine_prim_dev(B, 0, &pb);
    if (st != SHLL_OK) return st;
    return compare_dev(A, A->dim, pa, A->plane, pb, B->plane, A->inst_cells[0], cell_volume(ctx->pb), out);
}

extern "C" shll_status shll_compare_host(shll_ctx* ctx, const double* ref_prim_soa, shll_errors* out) {
    if (!ctx || !ref_prim_soa || !out) return fail(SHLL_EINVAL, "NULL argument");
    Engine* E = ctx->E;
    HIPCHECK(hipSetDevice(E->device));
    const int64_t nc = E->inst_cells[0];
    const double* pa = nullptr;
    shll_status st = engine_prim_dev(E, 0, &pa);
    if (st != SHLL_OK) return st;
    if (!E->d_cons) HIPCHECK(hipMalloc(&E->d_cons, (size_t)E->plane * E->nv * sizeof(double)));
    for (int v = 0; v < E->nv; ++v)
        HIPCHECK(hipMemcpyAsync(E->d_cons + (size_t)v * E->plane, ref_prim_soa + (size_t)v * nc, nc * sizeof(double),
                                hipMemcpyHostToDevice, E->stream));
    HIPCHECK(hipStreamSynchronize(E->stream));
    return compare_dev(E, E->dim, pa, E->plane, E->d_cons, E->plane, nc, cell_volume(ctx->pb), out);
}

static shll_status set_launch(Engine* E, int32_t C, int32_t k, int32_t wpb, int32_t graph) {
    if (E->dim == 1 && !step1d_supported(C, wpb))
        return fail(SHLL_EINVAL, "unsupported launch shape C=%d waves/block=%d", C, wpb);
    if (E->dim == 2 && wpb != 1 && wpb != 4) return fail(SHLL_EINVAL, "waves/block must be 1 or 4");
    if (E->dim == 2 && (C < 0 || C > kMaxRows2d))
        return fail(SHLL_EINVAL, "2D rows per item (cells_per_lane) must be 0-%d (got %d)", kMaxRows2d, C);
    if (k < 1 || graph < 0) return fail(SHLL_EINVAL, "steps per launch must be >= 1, graph steps >= 0");
    // a graph is captured when the shape is set (single instances) or first replayed: bound
    // its size
    if (graph / k > kMaxGraphLaunches)
        return fail(SHLL_EINVAL, "graph steps / steps per launch must be at most %d (got %d / %d)", kMaxGraphLaunches,
                    graph, k);
    if (E->dim == 2 && !step2d_supported(E->prec, E->order, k))
        return fail(SHLL_EINVAL, "2D steps per launch: 1, 2, 3, 4 or 8 in first order; 1, or 2 in the "
                                 "FP32-compensated second order (got %d)", k);
    if (E->batch && E->launches_done != 0)
        return fail(SHLL_ESTATE, "a batch's launch shape can only change before it advances (reset first)");
    Launch old = E->L;
    E->L.C = E->dim == 1 ? C : (C <= 1 ? 1 : C);
    E->L.k = k;
    if (E->dim == 2) E->L.rows = C <= 1 ? rows2d(k) : C;
    E->L.wpb = wpb;
    E->L.graph = graph;
    E->items_valid = false;
    shll_status st = engine_build_items(E);
    if (st != SHLL_OK) {
        E->L = old;
        E->items_valid = false;
        (void)engine_build_items(E);
        return st;
    }
    // a single instance's graphs depend on the launch shape only: capture them now
    const int g = E->batch ? 0 : single_graph_launches(E, launch_k(E));
    for (int p = 0; p < 2 && g > 0; ++p)
        if ((st = single_graph(E, p, g)) != SHLL_OK) return st;
    return SHLL_OK;
}

extern "C" shll_status shll_set_launch(shll_ctx* ctx, int32_t C, int32_t k, int32_t wpb, int32_t graph) {
    if (!ctx) return fail(SHLL_EINVAL, "ctx is NULL");
    HIPCHECK(hipSetDevice(ctx->E->device));
    return set_launch(ctx->E, C, k, wpb, graph);
}

extern "C" shll_status shll_get_launch(const shll_ctx* ctx, int32_t out[4]) {
    if (!ctx || !out) return fail(SHLL_EINVAL, "NULL argument");
    out[0] = ctx->E->L.C;
    out[1] = ctx->E->L.k;
    out[2] = ctx->E->L.wpb;
    out[3] = ctx->E->L.graph;
    return SHLL_OK;
}

extern "C" shll_status shll_set_timing(shll_ctx* ctx, int32_t on) {
    if (!ctx) return fail(SHLL_EINVAL, "ctx is NULL");
    ctx->E->timing = on != 0;
    return SHLL_OK;
}

extern "C" shll_status shll_last_timing(const shll_ctx* ctx, double* avg, int64_t* launches, double* total) {
    if (!ctx) return fail(SHLL_EINVAL, "ctx is NULL");
    if (avg) *avg = ctx->E->last_avg_ms;
    if (launches) *launches = ctx->E->last_launches;
    if (total) *total = ctx->E->last_total_ms;
    return SHLL_OK;
}

// ---------------------------------------------------------------------------
// autotuner: time candidate launch shapes from a saved state, keep the fastest
// ---------------------------------------------------------------------------
static std::string tune_key(const Engine* E, int64_t shape_class) {
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, E->device) == hipSuccess) arch = prop.gcnArchName;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s dim%d order%d%s prec%d batch%d class%lld", arch.c_str(), E->dim, E->order,
                  E->lim == SHLL_LIM_MC ? "mc" : "", E->prec, E->batch ? 1 : 0, (long long)shape_class);
    return buf;
}

static bool tune_cache_lookup(const std::string& key, Launch& L) {
    const char* path = std::getenv("SHLL_TUNE_CACHE");
    if (!path) return false;
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        char k[300];
        int C, kk, w, g;
        char* bar = std::strchr(line, '|');
        if (!bar) continue;
        size_t klen = (size_t)(bar - line);
        if (klen >= sizeof(k)) continue;
        std::memcpy(k, line, klen);
        k[klen] = 0;
        if (key == k && std::sscanf(bar + 1, "%d %d %d %d", &C, &kk, &w, &g) == 4) {
            L.C = C; L.k = kk; L.wpb = w; L.graph = g;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

static void tune_cache_store(const std::string& key, const Launch& L) {
    const char* path = std::getenv("SHLL_TUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s|%d %d %d %d\n", key.c_str(), L.C, L.k, L.wpb, L.graph);
    std::fclose(f);
}

// 2D work items of every instance at k steps per launch and `rows` owned rows per item
// (the counts tile2d produces)
static int64_t items2d(const Engine* E, int k, int rows) {
    const int64_t H = (int64_t)E->order * k;
    int64_t n = 0;
    for (const Inst& in : E->h_inst) {
        int64_t strips = 1;
        if (in.ny > 64 - H) strips += (in.ny - (64 - H) + (64 - 2 * H) - 1) / (64 - 2 * H);
        n += strips * ((in.nx + rows - 1) / rows);
    }
    return n;
}

static int64_t simd_count(const Engine* E) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, E->device) != hipSuccess) return 1024;
    return 4LL * prop.multiProcessorCount;
}

static shll_status engine_autotune(Engine* E, int32_t trial_steps, int64_t shape_class,
                                   shll_status (*advance)(Engine*, int64_t)) {
    const std::string key = tune_key(E, shape_class);
    Launch cached;
    if (tune_cache_lookup(key, cached)) return set_launch(E, cached.C, cached.k, cached.wpb, cached.graph);
    shll_status st = engine_snapshot(E, true);
    if (st != SHLL_OK) return st;
    const Launch start = E->L;
    std::vector<Launch> cands;
    if (E->dim == 1) {
        for (int C : {1, 2, 4, 8})
            for (int w : {1, 4}) {
                std::vector<int> ks = {1, 2, 4, 8, 16, 32, 64};
                // a single instance that fits one segment or cooperative block has no halo, so
                // any number of steps fits one launch: try long launches too
                const int XW = step1d_items_per_block(C, w) == 1 && w > 1 ? w : 1;
                if (!E->batch && E->inst_cells[0] + 2 <= 64LL * C * XW)
                    for (int k : {256, 1024, 4096}) ks.push_back(k);
                for (int k : ks) {
                    // batches: fewer than 4 steps per launch only multiplies the launches
                    if (E->batch && k < 4) continue;
                    // single instances, whose launches are short, also try long graphs and
                    // direct launches (kernels of ~10 us outlast the host's launch cost, and a
                    // graph's nodes run with gaps between them: the 65536-cell tube ran 8%
                    // faster launched directly than in 128-step graphs)
                    std::vector<int> graphs = {std::max(128, 2 * k)};
                    if (!E->batch) {
                        if (graphs[0] < 2048) graphs.push_back(2048);
                        graphs.push_back(0);
                    }
                    for (int g : graphs) {
                        Launch L = start;
                        L.C = C; L.k = k; L.wpb = w;
                        L.graph = g;
                        cands.push_back(L);
                    }
                }
            }
    } else {
        for (int k : {1, 2, 3, 4, 8})
            for (int w : {1, 4}) {
                if (!step2d_supported(E->prec, E->order, k)) continue;
                // items a little longer or shorter than the default change how many rounds
                // of resident wavefronts a launch takes, whose last round may run part-empty;
                // much shorter items when the default ones give fewer than 4 wavefronts per
                // SIMD: each wavefront walks its rows one after another, so a launch of few
                // items is latency-bound, and more items trade halo rows for parallelism
                const int rd = rows2d(k);
                std::vector<int> rows = {1, rd * 3 / 4, rd * 5 / 4, rd * 3 / 2};
                if (items2d(E, k, rd) < 4 * simd_count(E))
                    for (int r : {rd / 2, rd * 3 / 8, rd / 4, rd * 3 / 16, rd / 8})
                        if (r >= 2) rows.push_back(r);
                std::vector<int> graphs = {std::max(64, 2 * k)};
                if (!E->batch) {
                    graphs.push_back(1024);
                    graphs.push_back(0);
                }
                for (int r : rows)
                    for (int g : graphs) {
                        Launch L = start;
                        L.C = r;
                        L.k = k;
                        L.wpb = w;
                        L.graph = g;
                        cands.push_back(L);
                    }
            }
    }
    double best = 1e300;
    Launch bestL = start;
    const bool was_timing = E->timing;
    E->timing = false;
    for (const Launch& L : cands) {
        engine_rewind(E);
        if (set_launch(E, L.C, L.k, L.wpb, L.graph) != SHLL_OK) continue;
        if ((st = engine_snapshot(E, false)) != SHLL_OK) break;
        engine_rewind(E);
        // warm-up (also captures graphs: a single instance runs two chunks of the candidate's
        // graph, so the timed trial never includes a capture), then the timed trial
        const int64_t warm = std::max<int64_t>(std::max<int64_t>(1, trial_steps / 4), E->batch ? 0 : 2LL * L.graph);
        if ((st = advance(E, warm)) != SHLL_OK) break;
        if ((st = engine_snapshot(E, false)) != SHLL_OK) break;
        engine_rewind(E);
        auto t0 = std::chrono::steady_clock::now();
        if ((st = advance(E, trial_steps)) != SHLL_OK) break;
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (dt < best) {
            best = dt;
            bestL = L;
        }
    }
    E->timing = was_timing;
    engine_rewind(E);
    shll_status st2 = set_launch(E, bestL.C, bestL.k, bestL.wpb, bestL.graph);
    shll_status st3 = engine_snapshot(E, false);
    engine_rewind(E);
    if (st != SHLL_OK) return st;
    if (st2 != SHLL_OK) return st2;
    if (st3 != SHLL_OK) return st3;
    tune_cache_store(key, bestL);
    return SHLL_OK;
}

static shll_status single_advance(Engine* E, int64_t nsteps) { return single_steps(E, nsteps); }

extern "C" shll_status shll_autotune(shll_ctx* ctx, int32_t trial_steps) {
    if (!ctx || trial_steps < 1) return fail(SHLL_EINVAL, "bad argument");
    HIPCHECK(hipSetDevice(ctx->E->device));
    int64_t cls = 0;
    for (int64_t c = ctx->E->inst_cells[0]; c > 1; c >>= 1) cls++;  // log2 of the cell count
    return engine_autotune(ctx->E, trial_steps, cls, single_advance);
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
struct shll_batch {
    std::vector<shll_problem> pbs;
    Engine* E;
    int64_t total_updates;
};

extern "C" shll_status shll_batch_init(const shll_problem* items, const double* const* prims, const int64_t* nsteps,
                                       int32_t n, shll_batch** out) {
    if (!out) return fail(SHLL_EINVAL, "out is NULL");
    *out = nullptr;
    if (!items || !prims || !nsteps || n < 1) return fail(SHLL_EINVAL, "bad batch arguments");
    shll_status st;
    for (int32_t i = 0; i < n; ++i) {
        if ((st = check_problem(&items[i])) != SHLL_OK) return st;
        if (items[i].dim != items[0].dim || items[i].order != items[0].order ||
            items[i].precision != items[0].precision || items[i].limiter != items[0].limiter)
            return fail(SHLL_EINVAL, "batch instances must share dim, order, precision and limiter");
        if (!prims[i]) return fail(SHLL_EINVAL, "prims[%d] is NULL", i);
        if (nsteps[i] < 0) return fail(SHLL_EINVAL, "nsteps[%d] < 0", i);
    }
    if ((st = ensure_device()) != SHLL_OK) return st;
    Engine* E = engine_new(items[0].dim, items[0].order, items[0].precision, items[0].limiter);
    E->batch = true;
    if ((st = engine_create_stream(E)) != SHLL_OK) { engine_free(E); return st; }
    int64_t off = 0, maxcells = 0, total = 0;
    for (int32_t i = 0; i < n; ++i) {
        Inst in;
        std::memset(&in, 0, sizeof(in));
        in.off = off;
        in.nsteps = nsteps[i];
        in.nx = (int32_t)items[i].nx;
        in.ny = (int32_t)items[i].ny;
        in.bc = items[i].bc;
        in.ke = ke_of(items[i].dim, items[i].order);
        in.k = make_consts(items[i].nx, items[i].ny, items[i].lx, items[i].dim == 2 ? items[i].ly : 1.0,
                           items[i].theta);
        const int64_t cells = items[i].nx * items[i].ny;
        E->h_inst.push_back(in);
        E->inst_cells.push_back(cells);
        off += (cells + 63) / 64 * 64;
        maxcells = std::max(maxcells, cells);
        total += cells * nsteps[i];
    }
    E->order_idx.resize(n);
    for (int32_t i = 0; i < n; ++i) E->order_idx[i] = i;
    std::stable_sort(E->order_idx.begin(), E->order_idx.end(),
                     [&](int64_t a, int64_t b) { return E->h_inst[a].nsteps > E->h_inst[b].nsteps; });
    default_launch(E, maxcells);
    if (E->dim == 1) {
        E->L.C = 2;
        E->L.k = 16;
        E->L.wpb = 4;
        E->L.graph = 128;
    }
    if ((st = engine_alloc(E, off)) != SHLL_OK) { engine_free(E); return st; }
    if ((st = engine_init_state(E, prims)) != SHLL_OK) { engine_free(E); return st; }
    if ((st = engine_snapshot(E, true)) != SHLL_OK) { engine_free(E); return st; }
    if ((st = engine_build_items(E)) != SHLL_OK) { engine_free(E); return st; }
    shll_batch* b = new shll_batch;
    b->pbs.assign(items, items + n);
    b->E = E;
    b->total_updates = total;
    *out = b;
    return SHLL_OK;
}

extern "C" void shll_batch_free(shll_batch* b) {
    if (!b) return;
    engine_free(b->E);
    delete b;
}

static shll_status batch_advance_steps(Engine* E, int64_t nsteps) { return batch_launches(E, nsteps); }

extern "C" shll_status shll_batch_advance(shll_batch* b, int64_t nsteps, int64_t* max_remaining) {
    if (!b || nsteps < 0) return fail(SHLL_EINVAL, "bad argument");
    Engine* E = b->E;
    HIPCHECK(hipSetDevice(E->device));
    shll_status st = engine_build_items(E);
    if (st != SHLL_OK) return st;
    st = batch_advance_steps(E, nsteps);
    if (st != SHLL_OK) return st;
    if (max_remaining) {
        *max_remaining = std::max<int64_t>(0, batch_total_steps(E) - E->steps_done);
    }
    return SHLL_OK;
}

extern "C" shll_status shll_batch_run(shll_batch* b) {
    if (!b) return fail(SHLL_EINVAL, "batch is NULL");
    const int64_t total = b->E->h_inst[b->E->order_idx[0]].nsteps;
    return shll_batch_advance(b, total, nullptr);
}

extern "C" shll_status shll_batch_get_primitives(shll_batch* b, int32_t k, double* out) {
    if (!b || !out || k < 0 || k >= (int32_t)b->pbs.size()) return fail(SHLL_EINVAL, "bad argument");
    HIPCHECK(hipSetDevice(b->E->device));
    return engine_get(b->E, k, out, nullptr);
}

extern "C" shll_status shll_batch_get_conserved(shll_batch* b, int32_t k, double* out) {
    if (!b || !out || k < 0 || k >= (int32_t)b->pbs.size()) return fail(SHLL_EINVAL, "bad argument");
    HIPCHECK(hipSetDevice(b->E->device));
    return engine_get(b->E, k, nullptr, out);
}

extern "C" shll_status shll_batch_conserved_sums(shll_batch* b, double* out) {
    if (!b || !out) return fail(SHLL_EINVAL, "bad argument");
    HIPCHECK(hipSetDevice(b->E->device));
    std::vector<double> s;
    shll_status st = engine_sums(b->E, s);
    if (st != SHLL_OK) return st;
    for (size_t i = 0; i < b->pbs.size(); ++i) {
        const double vol = cell_volume(b->pbs[i]);
        for (int v = 0; v < 4; ++v) out[i * 4 + v] = v < b->E->nv ? s[i * 5 + v] * vol : 0.0;
    }
    return SHLL_OK;
}

extern "C" shll_status shll_batch_check_state(shll_batch* b, int64_t* bad_cells, int32_t* bad_instances) {
    if (!b) return fail(SHLL_EINVAL, "batch is NULL");
    HIPCHECK(hipSetDevice(b->E->device));
    std::vector<double> s;
    shll_status st = engine_sums(b->E, s);
    if (st != SHLL_OK) return st;
    int32_t nbad = 0;
    for (size_t i = 0; i < b->pbs.size(); ++i) {
        const int64_t bad = (int64_t)s[i * 5 + 4];
        if (bad_cells) bad_cells[i] = bad;
        nbad += bad > 0;
    }
    if (bad_instances) *bad_instances = nbad;
    return SHLL_OK;
}

extern "C" shll_status shll_batch_compare(shll_batch* b, int32_t k, shll_batch* ref, int32_t kref,
                                          shll_errors* out) {
    if (!b || !ref || !out || k < 0 || k >= (int32_t)b->pbs.size() || kref < 0 || kref >= (int32_t)ref->pbs.size())
        return fail(SHLL_EINVAL, "bad argument");
    const shll_problem &pa = b->pbs[k], &pr = ref->pbs[kref];
    if (pa.dim != pr.dim || pa.nx != pr.nx || pa.ny != pr.ny) return fail(SHLL_EINVAL, "instances differ in shape");
    if (b->E->device != ref->E->device) return fail(SHLL_EINVAL, "batches on different devices");
    HIPCHECK(hipSetDevice(b->E->device));
    const double *x = nullptr, *y = nullptr;
    shll_status st = engine_prim_dev(b->E, k, &x);
    if (st == SHLL_OK) st = engine_prim_dev(ref->E, kref, &y);
    if (st != SHLL_OK) return st;
    return compare_dev(b->E, pa.dim, x, b->E->plane, y, ref->E->plane, b->E->inst_cells[k], cell_volume(pa), out);
}

extern "C" shll_status shll_batch_set_launch(shll_batch* b, int32_t C, int32_t k, int32_t wpb, int32_t graph) {
    if (!b) return fail(SHLL_EINVAL, "batch is NULL");
    HIPCHECK(hipSetDevice(b->E->device));
    return set_launch(b->E, C, k, wpb, graph);
}

extern "C" shll_status shll_batch_get_launch(const shll_batch* b, int32_t out[4]) {
    if (!b || !out) return fail(SHLL_EINVAL, "NULL argument");
    out[0] = b->E->L.C;
    out[1] = b->E->L.k;
    out[2] = b->E->L.wpb;
    out[3] = b->E->L.graph;
    return SHLL_OK;
}

extern "C" shll_status shll_batch_reset(shll_batch* b) {
    if (!b) return fail(SHLL_EINVAL, "batch is NULL");
    HIPCHECK(hipSetDevice(b->E->device));
    shll_status st = engine_snapshot(b->E, false);
    engine_rewind(b->E);
    return st;
}

extern "C" shll_status shll_batch_autotune(shll_batch* b, int32_t trial_steps) {
    if (!b || trial_steps < 1) return fail(SHLL_EINVAL, "bad argument");
    HIPCHECK(hipSetDevice(b->E->device));
    if (b->E->launches_done != 0) return fail(SHLL_ESTATE, "autotune a batch before it advances (reset first)");
    int64_t cls = 0;
    for (int64_t c = b->total_updates; c > 1; c >>= 2) cls++;
    return engine_autotune(b->E, trial_steps, cls, batch_advance_steps);
}

extern "C" shll_status shll_batch_set_timing(shll_batch* b, int32_t on) {
    if (!b) return fail(SHLL_EINVAL, "batch is NULL");
    b->E->timing = on != 0;
    return SHLL_OK;
}

extern "C" shll_status shll_batch_last_timing(const shll_batch* b, double* avg, int64_t* launches, double* total) {
    if (!b) return fail(SHLL_EINVAL, "batch is NULL");
    if (avg) *avg = b->E->last_avg_ms;
    if (launches) *launches = b->E->last_launches;
    if (total) *total = b->E->last_total_ms;
    return SHLL_OK;
}

extern "C" int64_t shll_batch_total_cell_updates(const shll_batch* b) { return b ? b->total_updates : 0; }

// ---------------------------------------------------------------------------
// initial data
// ---------------------------------------------------------------------------
extern "C" void shll_ic_tube(int64_t n, double x0, double rhoL, double uL, double TL, double rhoR, double uR,
                             double TR, double* prim) {
    for (int64_t i = 0; i < n; ++i) {
        const bool left = (double)i < x0 * (double)n;  // base_shll.c:55: i < 0.5*N
        prim[i] = (double)(float)(left ? rhoL : rhoR);
        prim[n + i] = (double)(float)(left ? uL : uR);
        prim[2 * n + i] = (double)(float)(left ? TL : TR);
    }
}

extern "C" void shll_ic_implosion(int64_t nx, int64_t ny, double* prim) {
    const int64_t n = nx * ny;
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j) {
            const int64_t c = i * ny + j;
            const bool in = (i > 0.2 * nx) && (i < 0.8 * nx) && (j > 0.2 * ny) && (j < 0.8 * ny);
            prim[c] = in ? 1.0 : 10.0;
            prim[n + c] = 0.0;
            prim[2 * n + c] = 0.0;
            prim[3 * n + c] = 1.0;
        }
}

extern "C" void shll_ic_four_shocks(int64_t nx, int64_t ny, double* prim) {
    const int64_t n = nx * ny;
    const float R = 1.0f;
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j) {
            const int64_t c = i * ny + j;
            float r, vx, vy, T;
            if ((i < 0.75 * nx) && (j < 0.75 * ny)) {
                r = 0.138f; vx = 1.206f; vy = 1.206f; T = (float)(0.029 / (double)(r * R));
            } else if ((i > 0.75 * nx) && (j < 0.75 * ny)) {
                r = 0.5323f; vx = 0.0f; vy = 1.206f; T = (float)(0.3 / (double)(r * R));
            } else if ((i < 0.75 * nx) && (j > 0.75 * ny)) {
                r = 0.5323f; vx = 1.206f; vy = 0.0f; T = (float)(0.3 / (double)(r * R));
            } else {
                r = 1.5f; vx = 0.0f; vy = 0.0f; T = (float)(1.5 / (double)(r * R));
            }
            prim[c] = r; prim[n + c] = vx; prim[2 * n + c] = vy; prim[3 * n + c] = T;
        }
}

// base-omp/2nd_order_base_shll.c:149-167 ("Configuration 6"): four quadrants around the
// centre, split by the reference's double tests i < 0.5*NX, i > 0.5*NX, j < 0.5*NY,
// j > 0.5*NY (cells on a centre line take the last region); T = 1/(rho*R) rounded to float.
extern "C" void shll_ic_four_contacts(int64_t nx, int64_t ny, double* prim) {
    struct Region {
        float rho, u, v;
    };
    static const Region lower_left = {1.0f, -0.75f, 0.5f}, lower_right = {3.0f, -0.75f, -0.5f},
                        upper_left = {2.0f, 0.75f, 0.5f}, rest = {1.0f, 0.75f, -0.5f};
    const int64_t n = nx * ny;
    const float R = 1.0f;
    for (int64_t i = 0; i < nx; ++i) {
        const bool ilo = i < 0.5 * nx, ihi = i > 0.5 * nx;
        for (int64_t j = 0; j < ny; ++j) {
            const bool jlo = j < 0.5 * ny, jhi = j > 0.5 * ny;
            const Region& g = (ilo && jlo) ? lower_left : (ihi && jlo) ? lower_right : (ilo && jhi) ? upper_left : rest;
            const int64_t c = i * ny + j;
            prim[c] = g.rho;
            prim[n + c] = g.u;
            prim[2 * n + c] = g.v;
            prim[3 * n + c] = (float)(1.0 / (double)(g.rho * R));
        }
    }
}

extern "C" void shll_ic_circle(int64_t nx, int64_t ny, double lx, double ly, double cx, double cy, double r,
                               double rho_in, double T_in, double rho_out, double T_out, double* prim) {
    const int64_t n = nx * ny;
    const double dx = lx / nx, dy = ly / ny;
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j) {
            const int64_t c = i * ny + j;
            const double x = (i + 0.5) * dx - cx, y = (j + 0.5) * dy - cy;
            const bool in = x * x + y * y < r * r;
            prim[c] = (double)(float)(in ? rho_in : rho_out);
            prim[n + c] = 0.0;
            prim[2 * n + c] = 0.0;
            prim[3 * n + c] = (double)(float)(in ? T_in : T_out);
        }
}

// SplitMix64 (Steele, Lea, Flood, OOPSLA 2014): a counter-based stream per (seed, k, q)
static uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unit(uint64_t seed, int64_t k, int q) {
    const uint64_t h = splitmix64(splitmix64(seed) ^ ((uint64_t)k * 64 + (uint64_t)q));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
static double to_float(double x) { return (double)(float)x; }

extern "C" void shll_gen_tube(uint64_t seed, int64_t k, int64_t nmin, int64_t nmax, shll_tube_spec* out) {
    const double a = std::log((double)nmin), b = std::log((double)nmax);
    int64_t n = (int64_t)std::llround(std::exp(a + unit(seed, k, 0) * (b - a)));
    out->n = std::min(nmax, std::max(nmin, n));
    out->rhoR = 1.0;
    out->TR = 1.0;
    out->rhoL = to_float(2.0 + 18.0 * unit(seed, k, 1));
    out->TL = to_float(0.5 + 1.5 * unit(seed, k, 2));
    out->x0 = to_float(0.3 + 0.4 * unit(seed, k, 3));
}

extern "C" void shll_gen_domain(uint64_t seed, int64_t k, int64_t smin, int64_t smax, shll_domain_spec* out) {
    std::memset(out, 0, sizeof(*out));
    const double a = std::log((double)smin), b = std::log((double)smax);
    out->nx = std::min(smax, std::max(smin, (int64_t)std::llround(std::exp(a + unit(seed, k, 0) * (b - a)))));
    out->ny = std::min(smax, std::max(smin, (int64_t)std::llround(std::exp(a + unit(seed, k, 1) * (b - a)))));
    const double h = 1.0 / (double)std::max(out->nx, out->ny);
    out->lx = (double)out->nx * h;
    out->ly = (double)out->ny * h;
    out->ncircles = 1 + std::min(3, (int)(4.0 * unit(seed, k, 2)));
    const double side = std::min(out->lx, out->ly);
    for (int c = 0; c < out->ncircles; ++c) {
        out->rho[c] = to_float(2.0 + 18.0 * unit(seed, k, 8 + 5 * c));
        out->T[c] = to_float(0.5 + 1.5 * unit(seed, k, 9 + 5 * c));
        out->r[c] = (0.05 + 0.25 * unit(seed, k, 10 + 5 * c)) * side;
        out->cx[c] = unit(seed, k, 11 + 5 * c) * out->lx;
        out->cy[c] = unit(seed, k, 12 + 5 * c) * out->ly;
    }
}

extern "C" void shll_ic_domain(const shll_domain_spec* d, double* prim) {
    const int64_t nx = d->nx, ny = d->ny, n = nx * ny;
    const double dx = d->lx / nx, dy = d->ly / ny;
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j) {
            const int64_t c = i * ny + j;
            double rho = 1.0, T = 1.0;
            const double x = (i + 0.5) * dx, y = (j + 0.5) * dy;
            for (int q = 0; q < d->ncircles; ++q) {
                const double ex = x - d->cx[q], ey = y - d->cy[q];
                if (ex * ex + ey * ey < d->r[q] * d->r[q]) {
                    rho = d->rho[q];
                    T = d->T[q];
                }
            }
            prim[c] = rho;
            prim[n + c] = 0.0;
            prim[2 * n + c] = 0.0;
            prim[3 * n + c] = T;
        }
}
