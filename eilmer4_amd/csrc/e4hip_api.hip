// e4hip_api.hip -- C ABI (include/e4hip.h) of the MI355X explicit
// gas-dynamic update.  Host bookkeeping of gas models and blocks, device
// buffers, and the launch sequence of gasdynamic_explicit_increment_with_fixed_grid
// (eilmer/simcore.d:2303-2952) and determine_time_step_size
// (eilmer/simcore.d:1347-1452).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/e4hip.h"
#include "aux_kernels.h"

namespace {

using e4::IdealGas;
using e4::L2R2W;
using e4::Prim;

using e4::TILE_X;
using e4::TILE_Y;

enum { GAS_IDEAL = 0, GAS_TP = 1 };

struct GasModelH {
    int type = GAS_IDEAL;
    IdealGas ideal{};
    double mMass = 0.0;
    int nsp = 1;
    e4::TPGas tp{};
    e4::TPGas* tp_dev = nullptr;     // device copy read by the kernels
    e4::GuessModel* guess_dev = nullptr;
    __bf16* guess_w = nullptr;       // W1 [64][32], W2 [64][64], W3 [64][64] in BF16
    double a_init = 0.0;             // mixture: sound speed of the constructed cell state
};

struct BlockH {
    bool alive = false;
    int gm = -1;
    int dims = 3, ni = 0, nj = 0, nk = 0;
    double dx = 0, dy = 0, dz = 0;
    long ncells = 0, qs = 0;
    int nq = 5;
    int nsp = 1;
    // U[0] is the current state; stage s writes U[s], and a step that ends
    // without bad cells swaps U[0] and U[stages] (simcore.d:2936-2941), so a
    // step that stops leaves U[0] as it was
    double* U[4] = {nullptr, nullptr, nullptr, nullptr};
    double* d[2] = {nullptr, nullptr};            // dUdt of stages 1 and 2
    int bc_type[6] = {0, 0, 0, 0, 0, 0};
    Prim bc_fs[6] = {};
    e4::BCTable* bc_dev = nullptr;                // device copy of bc_type, bc_fs and connections
    int nbr_blk[6] = {-1, -1, -1, -1, -1, -1};    // connected block per face (E4HIP_BC_CONNECTED)
    int nbr_face[6] = {0, 0, 0, 0, 0, 0};
    int nbr_orient[6] = {0, 0, 0, 0, 0, 0};
    // thermally perfect gas only: primitive arrays (stage s reads P[s-1] and
    // writes P[s]; swapped like U) and face fluxes
    double* P[4] = {nullptr, nullptr, nullptr, nullptr};
    double* F[3] = {nullptr, nullptr, nullptr};
    long fqs[3] = {0, 0, 0};
    unsigned int* retry = nullptr;                // faces the mixed form hands on (e4::TPArgs::retry)
    double bc_fs_tp[6][8 + e4::TP_NSP_MAX] = {};
    // stored ghost cells per face (x_order-1 extrapolation of a mixture, static
    // profiles): [8+nsp][2 * face cells]; gS the profile's shock indicators
    double* gx[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    unsigned char* gS[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double* axi = nullptr;                        // 2D axisymmetric geometry (e4::axi_geom layout)
    double* src = nullptr;                        // user-defined source terms [nq][qs]
    // shock indicators per cell (adaptive_hanel_ausmdv): S[0] this stage's
    // detection, S[1] the previous one
    unsigned char* S[2] = {nullptr, nullptr};
    // a block built from its grid (e4hip_block_create_geometry): cell geometry
    // [4][qs] (volume, i/j/k lengths) and face geometry [10][fqs[d]] (area, n,
    // t1, t2); its stages run on the face kernels of gen_kernels.h with a
    // flow-state array P, whatever the gas
    bool general = false;
    double* geo = nullptr;
    double* fgeo[3] = {nullptr, nullptr, nullptr};
    bool is_tp() const;
    bool has_P() const { return general || is_tp(); }
};

struct ProfEvent { int cls; hipEvent_t a, b; };

// Staging of the batched launches: a slot per (set, stage, dimensionality).
// Set 0 serves steps run eagerly; sets 1 and 2 belong to the two captured
// step graphs, whose copy nodes read their pinned host slots whenever they
// replay, so only the capture of that graph writes them.
constexpr int BATCH_SETS = 3;
constexpr int BATCH_SLOTS = BATCH_SETS * 3 * 2;
struct BatchBuffers {
    void* args_dev = nullptr;
    void* args_host = nullptr;
    void* map_dev = nullptr;
    void* map_host = nullptr;
    size_t args_slot = 0, map_slot = 0;
    // set 0: bytes of each slot whose device copy equals its host copy (0 after a regrow)
    size_t args_valid[BATCH_SLOTS] = {}, map_valid[BATCH_SLOTS] = {};
};

inline int batch_slot(int set, int stage, int dims) { return (set * 3 + (stage - 1)) * 2 + (dims - 2); }

// A step captured into a HIP graph: the key of the state it starts from
// (generation, blocks, dt, buffer pointers), the blocks' buffer pointers after
// it, and the batch staging set its copy nodes read.
struct StepGraph {
    std::vector<uintptr_t> key, post;
    hipGraphExec_t exec = nullptr;
    int set = 1;
    unsigned long long last_use = 0;
};

// Words before the block-stage counter slots: the abort flag of stage s at
// [32 (s-1)], each in a 128-byte line of its own, so the per-wave reads of
// the previous stage's flag never meet the atomics on the counters.
constexpr int ABORT_WORDS = 96;
constexpr int ABORT_STRIDE = 32;
// pointers per block in a step's snapshot (U[4], d[2], P[4], S[2])
constexpr int PTRS_PER_BLOCK = 12;

struct State {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned int* counters_raw = nullptr;      // ABORT_WORDS abort flags, then the counter slots
    unsigned int* abort = nullptr;             // = counters_raw: abort flag of stage s at [ABORT_STRIDE (s-1)]
    unsigned int* counters = nullptr;          // per-stage counters, e4::COUNTER_SLOT uints per block-stage slot
    unsigned int* counters_host = nullptr;     // pinned host copy the step reads the flags and counters back into
    unsigned int* step_abort = nullptr;        // abort flags of the step being enqueued
    unsigned long long* smax = nullptr;
    double* last_signal = nullptr;
    int counter_slots = 0;
    std::vector<GasModelH> gases;
    BatchBuffers batch;
    BatchBuffers fill;                         // staging of the batched x_order-1 ghost fill
    BatchBuffers tpbatch;                      // staging of the batched mixture launches
    int batch_min_blocks = 2;                  // batch the ideal blocks of a stage from this many on
    std::vector<BlockH> blocks;
    E4hipStepConfig cfg{E4HIP_FLUX_AUSMDV, E4HIP_UPDATE_PC, 2, 1, 1, 1, E4HIP_THERMO_RHOU, E4HIP_PRECISION_FP64, 0, 0,
                        0, 0.20, -0.30, 0, 0, 0.01, 1};
    bool profiling = false;
    std::vector<ProfEvent> prof_pending;
    double prof_ms[3] = {0, 0, 0};
    long prof_n[3] = {0, 0, 0};
    unsigned long long* stats = nullptr;               // e4::add_stat's copies (device)
    std::vector<int> last_bad;                          // [nblk][3] of the last step: invalid cells + failed decodes
    // graph replay of steps (use_graphs): up to two captured steps (a step
    // swaps U[0] and U[stages], so the starting states of consecutive steps
    // alternate between two keys), and the keys of the last two steps run
    // without a graph
    std::vector<StepGraph> graphs;
    std::vector<std::vector<uintptr_t>> recent_keys;
    unsigned long long step_counter = 0;
    unsigned long long generation = 0;         // bumped by every call that changes blocks, boundaries or config
    // k-chunk autotuner: chunk height of the stage kernel per block shape and
    // arithmetic (ni, nj, nk, precision, kernel flux, default reconstruction)
    std::vector<std::pair<std::vector<int>, int>> kchunk_tuned;
    // work queue of persistent batched stage launches (next tile, finished
    // workgroups; cleared by the last workgroup of each launch) and the GPU's
    // compute units
    unsigned int* queue = nullptr;
    int cus = 0;
    std::string err;
};

State g_state;

// ---- scratch of e4hip_determine_dt_noalloc, the entry point that allocates
// nothing: per block its largest CFL signal (bit pattern, atomic max) and its
// last cell's signal, then the reduced dt_allow and cfl_max.  One array per
// GPU, held until the process ends; its device address is taken once per
// device at e4hip_init (host code never passes the array's name to a kernel).
constexpr int NOALLOC_WORDS = 2 * E4HIP_NOALLOC_MAX_BLOCKS + 2;
static __device__ __align__(256) unsigned long long scratch[NOALLOC_WORDS];
constexpr int MAX_DEVICES = 64;
unsigned long long* scratch_dev[MAX_DEVICES] = {};

__global__ void noalloc_clear_kernel(unsigned long long* s, int n)
{
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) s[t] = 0ull;
}

// The blocks' (dt_allow, cfl_max) from their slots, reduced in block order as
// e4hip_determine_dt does on the host (block_dt, fluidblock.d:776-862,
// simcore.d:1385-1394), into out[0], out[1].
__global__ void noalloc_dt_reduce_kernel(const unsigned long long* s, int nblk, double dt_current, double cfl_value,
                                         int check_cfl, double cfl_allow, double* out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double dt_allow_all = 1.7976931348623157e308, cfl_max_all = 0.0;
    for (int n = 0; n < nblk; ++n) {
        const double smax = __longlong_as_double((long long)s[2 * n]);
        const double last = __longlong_as_double((long long)s[2 * n + 1]);
        double dt_allow = cfl_value / smax;
        double cfl_max = dt_current * smax;
        if (check_cfl && (cfl_max < 0.0 || cfl_max > cfl_allow)) {
            cfl_max = 0.5 * cfl_allow;
            dt_allow = cfl_max / last;
        }
        dt_allow_all = fmin(dt_allow_all, dt_allow);
        cfl_max_all = fmax(cfl_max_all, cfl_max);
    }
    out[0] = dt_allow_all;
    out[1] = cfl_max_all;
}

bool BlockH::is_tp() const { return g_state.gases[gm].type == GAS_TP; }

void graphs_drop()
{
    for (auto& g : g_state.graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g_state.graphs.clear();
    g_state.recent_keys.clear();
}

// Any call that changes what a step's kernels are given drops the captured steps.
void graph_invalidate()
{
    ++g_state.generation;
    graphs_drop();
}

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_state.err = buf;
    return code;
}

#define HIP_TRY(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(E4HIP_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

BlockH* get_block(int h)
{
    if (h < 0 || h >= (int)g_state.blocks.size() || !g_state.blocks[h].alive) return nullptr;
    return &g_state.blocks[h];
}

int nstages(int scheme)
{
    switch (scheme) {
    case E4HIP_UPDATE_EULER: return 1;
    case E4HIP_UPDATE_PC: case E4HIP_UPDATE_MIDPOINT: return 2;
    default: return 3;
    }
}

// l2r2_prepare (eilmer/onedinterp.d:118-133) for a uniform direction of cell length len
L2R2W l2r2_weights(double len)
{
    double lenL1 = len, lenL0 = len, lenR0 = len, lenR1 = len;
    L2R2W w;
    w.lenL0 = lenL0;
    w.lenR0 = lenR0;
    w.aL0 = 0.5 * lenL0 / (lenL1 + 2.0 * lenL0 + lenR0);
    w.aR0 = 0.5 * lenR0 / (lenL0 + 2.0 * lenR0 + lenR1);
    w.two_over_lenL0_plus_lenL1 = 2.0 / (lenL0 + lenL1);
    w.two_over_lenR0_plus_lenL0 = 2.0 / (lenR0 + lenL0);
    w.two_over_lenR1_plus_lenR0 = 2.0 / (lenR1 + lenR0);
    w.two_lenL0_plus_lenL1 = (2.0 * lenL0 + lenL1);
    w.two_lenR0_plus_lenR1 = (2.0 * lenR0 + lenR1);
    return w;
}

// The four distinct numbers of l2r2_weights(len) that the stage kernel's
// per-cell reconstruction uses; the equalities hold because every length is len.
bool l2r2_uniform(double len, e4::L2R2U& u)
{
    L2R2W w = l2r2_weights(len);
    u.c = w.two_over_lenL0_plus_lenL1;
    u.a = w.aL0;
    u.len = w.lenR0;
    u.len3 = w.two_lenL0_plus_lenL1;
    return w.two_over_lenR0_plus_lenL0 == u.c && w.two_over_lenR1_plus_lenR0 == u.c && w.aR0 == u.a &&
           w.lenL0 == u.len && w.two_lenR0_plus_lenR1 == u.len3;
}

int kernel_flux(int fc);

std::vector<int> tune_key(const BlockH& b)
{
    const E4hipStepConfig& c = g_state.cfg;
    const int def = c.interpolation_order > 1 && c.apply_limiter && c.extrema_clipping;
    return {b.ni, b.nj, b.nk, c.precision_mode, kernel_flux(c.flux_calculator), def};
}

int kchunk_for(const BlockH& b)
{
    const char* env = getenv("E4HIP_KCHUNK");
    if (env && atoi(env) > 0) return atoi(env);
    const std::vector<int> key = tune_key(b);
    for (const auto& t : g_state.kchunk_tuned)
        if (t.first == key) return t.second;
    long nxy = (long)((b.ni + TILE_X - 1) / TILE_X) * ((b.nj + TILE_Y - 1) / TILE_Y);
    // aim for >= 2048 workgroups (8 per CU) while marching at least 8 planes
    long chunks = (2048 + nxy - 1) / nxy;
    long kc = (b.nk + chunks - 1) / chunks;
    if (kc < 8) kc = 8;
    if (kc > b.nk) kc = b.nk;
    return (int)kc;
}

int ensure_counters(int slots)
{
    if (slots <= g_state.counter_slots) return E4HIP_OK;
    // captured steps hold the old counter addresses
    graph_invalidate();
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    if (g_state.counters_raw) HIP_TRY(hipFree(g_state.counters_raw));
    if (g_state.counters_host) HIP_TRY(hipHostFree(g_state.counters_host));
    g_state.counters_raw = g_state.abort = g_state.counters = nullptr;
    g_state.counters_host = nullptr;
    g_state.counter_slots = 0;
    const size_t n = ABORT_WORDS + (size_t)e4::COUNTER_SLOT * slots;
    HIP_TRY(hipMalloc(&g_state.counters_raw, sizeof(unsigned int) * n));
    HIP_TRY(hipHostMalloc(&g_state.counters_host, sizeof(unsigned int) * n));
    g_state.abort = g_state.counters_raw;
    g_state.counters = g_state.counters_raw + ABORT_WORDS;
    g_state.counter_slots = slots;
    return E4HIP_OK;
}

// The solver statistics: the sums of counters ka and kb over e4::add_stat's
// copies (after the stream drains); reset clears both in every copy.
int read_stats(int ka, int kb, long long* a, long long* b, int reset)
{
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    const int n = e4::STAT_LANES * e4::STAT_STRIDE;
    std::vector<unsigned long long> h(n);
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    HIP_TRY(hipMemcpy(h.data(), g_state.stats, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
    unsigned long long sa = 0, sb = 0;
    for (int l = 0; l < e4::STAT_LANES; ++l) {
        sa += h[l * e4::STAT_STRIDE + ka];
        sb += h[l * e4::STAT_STRIDE + kb];
        if (reset) { h[l * e4::STAT_STRIDE + ka] = 0; h[l * e4::STAT_STRIDE + kb] = 0; }
    }
    if (a) *a = (long long)sa;
    if (b) *b = (long long)sb;
    if (reset) HIP_TRY(hipMemcpy(g_state.stats, h.data(), sizeof(unsigned long long) * n, hipMemcpyHostToDevice));
    return E4HIP_OK;
}

void prof_begin(hipEvent_t* a)
{
    if (!g_state.profiling) return;
    hipEventCreate(a);
    hipEventRecord(*a, g_state.stream);
}

void prof_end(int cls, hipEvent_t a)
{
    if (!g_state.profiling) return;
    hipEvent_t b;
    hipEventCreate(&b);
    hipEventRecord(b, g_state.stream);
    g_state.prof_pending.push_back(ProfEvent{cls, a, b});
}

e4::Connection connection_of(const BlockH& b, int f)
{
    e4::Connection c;
    memset(&c, 0, sizeof(c));
    const BlockH* o = get_block(b.nbr_blk[f]);
    if (b.bc_type[f] != E4HIP_BC_CONNECTED || !o) return c;
    c.n[0] = o->ni; c.n[1] = o->nj; c.n[2] = (o->dims == 3) ? o->nk : 1;
    c.face = b.nbr_face[f]; c.orient = b.nbr_orient[f]; c.dims = b.dims; c.qs = o->qs;
    return c;
}

// The kernels' flux instantiation of a FluxCalculator value.
int kernel_flux(int fc)
{
    if (fc == E4HIP_FLUX_HLLC) return e4::FLUX_HLLC;
    if (fc == E4HIP_FLUX_HANEL || fc == E4HIP_FLUX_ADAPTIVE_HANEL_AUSMDV || fc == E4HIP_FLUX_EFM ||
        fc == E4HIP_FLUX_ADAPTIVE_EFM_AUSMDV)
        return e4::FLUX_SHOCK;
    if (fc == E4HIP_FLUX_AUSM_PLUS_UP || fc == E4HIP_FLUX_ROE) return e4::FLUX_OTHER;
    return e4::FLUX_AUSMDV;
}

// Whether the flux calculator switches the shock detector on
// (globalconfig.d:1942-1954; adaptive_hlle_roe is not available here).
bool flux_detects_shocks(int fc)
{
    return fc == E4HIP_FLUX_ADAPTIVE_HANEL_AUSMDV || fc == E4HIP_FLUX_ADAPTIVE_EFM_AUSMDV;
}

// Shock-indicator arguments of a block's stage (StageArgs or TPArgs).
template <class Args>
int shock_fields(const BlockH& b, Args& A)
{
    const E4hipStepConfig& cfg = g_state.cfg;
    A.S = b.S[0];
    A.Sprev = b.S[1];
    A.Snew = b.S[0];
    for (int f = 0; f < 6; ++f) {
        if (b.bc_type[f] != E4HIP_BC_CONNECTED) continue;
        const BlockH* o = get_block(b.nbr_blk[f]);
        if (!o) return fail(E4HIP_ERR_HANDLE, "face %d is connected to a destroyed block", f);
        A.nbrS[f] = o->S[1];
    }
    A.shear_tol = cfg.shear_tolerance;
    A.compression_tol = cfg.compression_tolerance;
    const int fc = cfg.flux_calculator;
    A.always_hanel = fc == E4HIP_FLUX_HANEL || fc == E4HIP_FLUX_EFM;
    A.shock_efm = fc == E4HIP_FLUX_EFM || fc == E4HIP_FLUX_ADAPTIVE_EFM_AUSMDV;
    A.other_roe = fc == E4HIP_FLUX_ROE;
    A.M_inf = cfg.M_inf;
    return E4HIP_OK;
}

// The coefficients of the stage updates of fvcell.d:1153-1393:
// stage 1 U1 = U0 + dt g1 d0; stage 2 U2 = Uold + dt (g1 d0 + g2 d1);
// stage 3 U3 = Uold + dt (g1 d0 + g2 d1 + g3 d2).
void stage_coefficients(int sch, int stage, double& g1, double& g2, double& g3)
{
    g1 = 1.0; g2 = 0.0; g3 = 0.0;
    if (stage == 1) {
        if (sch == E4HIP_UPDATE_MIDPOINT || sch == E4HIP_UPDATE_CLASSIC_RK3) g1 = 0.5;
        if (sch == E4HIP_UPDATE_DENMAN_RK3) g1 = 8.0 / 15.0;
    } else if (stage == 2) {
        g1 = 0.5; g2 = 0.5;
        switch (sch) {
        case E4HIP_UPDATE_MIDPOINT: g1 = 0.0; g2 = 1.0; break;
        case E4HIP_UPDATE_CLASSIC_RK3: g1 = -1.0; g2 = 2.0; break;
        case E4HIP_UPDATE_TVD_RK3: g1 = 0.25; g2 = 0.25; break;
        case E4HIP_UPDATE_DENMAN_RK3: g1 = -17.0 / 60.0; g2 = 5.0 / 12.0; break;
        default: break;
        }
    } else {
        g1 = 1.0 / 6.0; g2 = 1.0 / 6.0; g3 = 4.0 / 6.0;
        if (sch == E4HIP_UPDATE_CLASSIC_RK3) { g1 = 1.0 / 6.0; g2 = 4.0 / 6.0; g3 = 1.0 / 6.0; }
        if (sch == E4HIP_UPDATE_DENMAN_RK3) { g1 = 0.0; g2 = -5.0 / 12.0; g3 = 3.0 / 4.0; }
    }
}

// The bad-cell stop of the step being enqueued: stage s reads the flag of
// stage s-1 and sets its own (StageArgs and TPArgs).
template <class Args>
void abort_fields(int stage, Args& A)
{
    unsigned int* ab = g_state.step_abort;
    A.abort_prev = (ab && stage > 1) ? ab + ABORT_STRIDE * (stage - 2) : nullptr;
    A.abort_this = ab ? ab + ABORT_STRIDE * (stage - 1) : nullptr;
    A.max_invalid = (unsigned int)g_state.cfg.max_invalid_cells;
}

// The stage kernel's arguments for one block; kchunk <= 0 picks the
// single-launch chunk.
void boundary_slabs(const BlockH& b, int* lo, int* hi, int* off, int& noghost, bool stored_ghosts);

int stage_args_for(const BlockH& b, int stage, double dt, unsigned int* counters, int kchunk, e4::StageArgs& A,
                   dim3& grid)
{
    const E4hipStepConfig& cfg = g_state.cfg;
    const int sch = cfg.gasdynamic_update_scheme;
    const int ns = nstages(sch);
    memset(&A, 0, sizeof(A));
    A.qs = b.qs;
    A.ni = b.ni; A.nj = b.nj; A.nk = (b.dims == 3) ? b.nk : 1;
    A.kchunk = (b.dims == 3) ? (kchunk > 0 ? kchunk : kchunk_for(b)) : 1;
    A.stage = stage;
    A.high_order = cfg.interpolation_order > 1;
    A.apply_limiter = cfg.apply_limiter;
    A.extrema_clipping = cfg.extrema_clipping;
    A.counters = counters;
    A.gas = g_state.gases[b.gm].ideal;
    A.bc = b.bc_dev;
    for (int f = 0; f < 6; ++f) {
        if (b.bc_type[f] != E4HIP_BC_CONNECTED) continue;
        const BlockH* o = get_block(b.nbr_blk[f]);
        if (!o) return fail(E4HIP_ERR_HANDLE, "face %d is connected to a destroyed block", f);
        A.nbrU[f] = o->U[stage - 1];   // the stencil state of stage s is U[s-1] in every block
    }
    // geometry of the uniform Cartesian cell (the oracle uses the same expressions)
    if (b.dims == 3) {
        A.vol_inv = 1.0 / (b.dx * b.dy * b.dz);
        A.area[0] = b.dy * b.dz; A.area[1] = b.dz * b.dx; A.area[2] = b.dx * b.dy;
    } else {
        A.vol_inv = 1.0 / (b.dx * b.dy);
        A.area[0] = b.dy; A.area[1] = b.dx; A.area[2] = 0.0;
    }
    if (!l2r2_uniform(b.dx, A.wu[0]) || !l2r2_uniform(b.dy, A.wu[1]) || !l2r2_uniform(b.dz, A.wu[2]))
        return fail(E4HIP_ERR_ARG, "l2r2 weights of a uniform direction are not uniform");
    // stage coefficients (fvcell.d:1168-1180, 1254-1266, 1333-1342) and
    // buffers: stage s evaluates its fluxes from U[s-1] and writes U[s]
    double g1 = 1.0, g2 = 0.0, g3 = 0.0;
    stage_coefficients(sch, stage, g1, g2, g3);
    A.Us = b.U[stage - 1];
    A.Unew = b.U[stage];
    if (stage == 1) {
        A.Uold = b.U[0];
        A.dout = (ns > 1) ? b.d[0] : nullptr;
    } else if (stage == 2) {
        A.d0 = b.d[0];
        A.Uold = (sch == E4HIP_UPDATE_DENMAN_RK3) ? b.U[1] : b.U[0];
        A.dout = (ns > 2) ? b.d[1] : nullptr;
    } else {
        A.d0 = b.d[0]; A.d1 = b.d[1];
        A.Uold = (sch == E4HIP_UPDATE_DENMAN_RK3) ? b.U[2] : b.U[0];
        A.dout = nullptr;
    }
    A.dt = dt; A.g1 = g1; A.g2 = g2; A.g3 = g3;
    A.c1 = dt * g1;
    abort_fields(stage, A);
    boundary_slabs(b, A.fix_lo, A.fix_hi, A.fix_off, A.noghost, true);
    for (int f = 0; f < 6; ++f) { A.gx[f] = b.gx[f]; A.gS[f] = b.gS[f]; }
    A.axi = b.axi;
    A.src = b.src;
    const int src = shock_fields(b, A);
    if (src) return src;
    const int nk = A.nk;
    A.tiles_x = (b.ni + TILE_X - 1) / TILE_X;
    A.tiles_y = (b.nj + TILE_Y - 1) / TILE_Y;
    grid = dim3(A.tiles_x, A.tiles_y, (nk + A.kchunk - 1) / A.kchunk);
    return E4HIP_OK;
}

// Cells the boundary kernel finishes: the layers beside faces whose flux is
// not the interior code's (two for a wall without ghost data, one for a
// simple-outflow-flux face) and, with stored_ghosts (the ideal-gas stage
// kernel, which does not read stored ghost states), two beside faces whose
// ghosts are stored (static profile, x_order-1 extrapolation); slab by slab.
void boundary_slabs(const BlockH& b, int* lo, int* hi, int* off, int& noghost, bool stored_ghosts)
{
    static const int FM[3] = {E4HIP_FACE_WEST, E4HIP_FACE_SOUTH, E4HIP_FACE_BOTTOM};
    static const int FP[3] = {E4HIP_FACE_EAST, E4HIP_FACE_NORTH, E4HIP_FACE_TOP};
    const int n[3] = {b.ni, b.nj, (b.dims == 3) ? b.nk : 1};
    auto layers = [&](int f) {
        const int t = b.bc_type[f];
        if (t == E4HIP_BC_SLIP_WALL_1) return 2;
        if (t == E4HIP_BC_OUTFLOW_SIMPLE_FLUX) return 1;
        if (stored_ghosts && (t == E4HIP_BC_INFLOW_PROFILE || t == E4HIP_BC_EXTRAPOLATE_1)) return 2;
        return 0;
    };
    noghost = 0;
    for (int d = 0; d < 3; ++d) {
        const bool on = d < b.dims;
        lo[d] = on ? layers(FM[d]) : 0;
        hi[d] = on ? layers(FP[d]) : 0;
        if (on && b.bc_type[FM[d]] == E4HIP_BC_SLIP_WALL_1) noghost |= 1 << FM[d];
        if (on && b.bc_type[FP[d]] == E4HIP_BC_SLIP_WALL_1) noghost |= 1 << FP[d];
    }
    int total = 0;
    for (int s = 0; s < 6; ++s) {
        const int d = s >> 1;
        off[s] = total;
        total += ((s & 1) ? hi[d] : lo[d]) * n[(d + 1) % 3] * n[(d + 2) % 3];
    }
    off[6] = total;
}

void launch_boundary_cells(const e4::StageArgs& A, int dims, int flux)
{
    if (A.fix_off[6] <= 0) return;
    switch (g_state.cfg.precision_mode) {
    case E4HIP_PRECISION_FP64_EXACT: e4::launch_boundary_cells_exact(A, g_state.stream, dims, flux); break;
    case E4HIP_PRECISION_MIXED: e4::launch_boundary_cells_mixed(A, g_state.stream, dims, flux); break;
    default: e4::launch_boundary_cells_fast64(A, g_state.stream, dims, flux); break;
    }
}

int launch_stage(const BlockH& b, int stage, double dt, unsigned int* counters)
{
    const E4hipStepConfig& cfg = g_state.cfg;
    e4::StageArgs A;
    dim3 grid;
    int rc = stage_args_for(b, stage, dt, counters, 0, A, grid);
    if (rc) return rc;
    hipEvent_t ev = nullptr;
    prof_begin(&ev);
    const int flux = kernel_flux(cfg.flux_calculator);
    switch (cfg.precision_mode) {
    case E4HIP_PRECISION_FP64_EXACT: e4::launch_stage_exact(A, grid, g_state.stream, b.dims, flux); break;
    case E4HIP_PRECISION_MIXED: e4::launch_stage_mixed(A, grid, g_state.stream, b.dims, flux); break;
    default: e4::launch_stage_fast64(A, grid, g_state.stream, b.dims, flux); break;
    }
    launch_boundary_cells(A, b.dims, flux);
    prof_end(0, ev);
    HIP_TRY(hipGetLastError());
    return E4HIP_OK;
}

// The k-chunk autotuner.  For a uniform 3D ideal-gas block whose shape has
// not been tuned under the current arithmetic, the stage kernel of stage 1 is
// timed (HIP events, three launches after a warm-up) at candidate chunk
// heights -- the heights that split nk into 1, 2, 3, 4, 6, 8, 12, 16 or 32
// chunks of at least four planes -- and the fastest is kept for that shape.
// Stage 1 only reads U[0] and writes U[1], d[0] and counters that the step
// resets, and every chunk height gives the same bits, so tuning before a step
// changes its timing only.
std::vector<int> batch_tune_key(const std::vector<const BlockH*>& blks);
int launch_stage_batch(const std::vector<const BlockH*>& blks, const std::vector<unsigned int*>& counters,
                       int stage, double dt, int set, long force_chunks);

int autotune_blocks(const int* blks, int nblk, double dt)
{
    if (!g_state.cfg.autotune) return E4HIP_OK;
    const char* env = getenv("E4HIP_KCHUNK");
    if (env && atoi(env) > 0) return E4HIP_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(g_state.stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return E4HIP_OK;
    const int flux = kernel_flux(g_state.cfg.flux_calculator);
    for (int n = 0; n < nblk; ++n) {
        const BlockH& b = *get_block(blks[n]);
        if (b.dims != 3 || b.has_P() || b.general) continue;
        const std::vector<int> key = tune_key(b);
        bool known = false;
        for (const auto& t : g_state.kchunk_tuned) known = known || t.first == key;
        if (known) continue;
        std::vector<int> cand;
        for (int chunks : {1, 2, 3, 4, 6, 8, 12, 16, 32}) {
            const int kc = (b.nk + chunks - 1) / chunks;
            if (kc >= 4 && std::find(cand.begin(), cand.end(), kc) == cand.end()) cand.push_back(kc);
        }
        if (cand.empty()) cand.push_back(b.nk);
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        // each candidate: the scheme's stages in order (the later stages read
        // more arrays than the first), one warm-up pass and three timed ones;
        // two rounds over the candidates, each keeping its faster round
        const int ns = nstages(g_state.cfg.gasdynamic_update_scheme);
        std::vector<float> cand_ms(cand.size(), 1e30f);
        for (int round = 0; round < 2; ++round) {
            for (size_t ci = 0; ci < cand.size(); ++ci) {
                const int kc = cand[ci];
                e4::StageArgs A[3];
                dim3 grid;
                for (int st = 1; st <= ns; ++st) {
                    int rc = stage_args_for(b, st, dt, g_state.counters, kc, A[st - 1], grid);
                    if (rc) return rc;
                }
                auto launch = [&]() {
                    for (int st = 0; st < ns; ++st) {
                        switch (g_state.cfg.precision_mode) {
                        case E4HIP_PRECISION_FP64_EXACT: e4::launch_stage_exact(A[st], grid, g_state.stream, 3, flux); break;
                        case E4HIP_PRECISION_MIXED: e4::launch_stage_mixed(A[st], grid, g_state.stream, 3, flux); break;
                        default: e4::launch_stage_fast64(A[st], grid, g_state.stream, 3, flux); break;
                        }
                    }
                };
                launch();
                HIP_TRY(hipEventRecord(e0, g_state.stream));
                for (int r = 0; r < 3; ++r) launch();
                HIP_TRY(hipEventRecord(e1, g_state.stream));
                HIP_TRY(hipEventSynchronize(e1));
                HIP_TRY(hipGetLastError());
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                cand_ms[ci] = std::min(cand_ms[ci], ms);
            }
        }
        int best = cand[0];
        float best_ms = 1e30f;
        for (size_t ci = 0; ci < cand.size(); ++ci)
            if (cand_ms[ci] < best_ms) { best_ms = cand_ms[ci]; best = cand[ci]; }
        HIP_TRY(hipEventDestroy(e0));
        HIP_TRY(hipEventDestroy(e1));
        g_state.kchunk_tuned.push_back({key, best});
        graph_invalidate();
    }
    // the ideal 3D blocks of a batched launch: the number of chunks per block
    // over the whole launch (one table upload per candidate, into the eager
    // staging set's stage-1 slot)
    if (g_state.cfg.batch_blocks) {
        std::vector<const BlockH*> bl;
        std::vector<unsigned int*> cnt;
        for (int n = 0; n < nblk; ++n) {
            const BlockH& b = *get_block(blks[n]);
            if (b.dims == 3 && !b.has_P()) {
                bl.push_back(&b);
                cnt.push_back(g_state.counters);
            }
        }
        if ((int)bl.size() >= g_state.batch_min_blocks) {
            const std::vector<int> key = batch_tune_key(bl);
            bool known = false;
            for (const auto& t : g_state.kchunk_tuned) known = known || t.first == key;
            if (!known) {
                hipEvent_t e0, e1;
                HIP_TRY(hipEventCreate(&e0));
                HIP_TRY(hipEventCreate(&e1));
                // as for one block: the scheme's stages in order, two rounds
                const int ns = nstages(g_state.cfg.gasdynamic_update_scheme);
                const long cand[] = {1L, 2L, 3L, 4L, 6L, 8L, 12L, 16L, 32L};
                float cand_ms[sizeof(cand) / sizeof(cand[0])];
                for (float& m : cand_ms) m = 1e30f;
                for (int round = 0; round < 2; ++round) {
                    for (size_t ci = 0; ci < sizeof(cand) / sizeof(cand[0]); ++ci) {
                        auto pass = [&]() {
                            for (int st = 1; st <= ns; ++st) {
                                int rc = launch_stage_batch(bl, cnt, st, dt, 0, cand[ci]);
                                if (rc) return rc;
                            }
                            return (int)E4HIP_OK;
                        };
                        int rc = pass();
                        if (rc) return rc;
                        HIP_TRY(hipEventRecord(e0, g_state.stream));
                        for (int r = 0; r < 3; ++r) {
                            rc = pass();
                            if (rc) return rc;
                        }
                        HIP_TRY(hipEventRecord(e1, g_state.stream));
                        HIP_TRY(hipEventSynchronize(e1));
                        HIP_TRY(hipGetLastError());
                        float ms = 0.0f;
                        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                        cand_ms[ci] = std::min(cand_ms[ci], ms);
                    }
                }
                long best = 1;
                float best_ms = 1e30f;
                for (size_t ci = 0; ci < sizeof(cand) / sizeof(cand[0]); ++ci)
                    if (cand_ms[ci] < best_ms) { best_ms = cand_ms[ci]; best = cand[ci]; }
                HIP_TRY(hipEventDestroy(e0));
                HIP_TRY(hipEventDestroy(e1));
                g_state.kchunk_tuned.push_back({key, (int)best});
                graph_invalidate();
            }
        }
    }
    return E4HIP_OK;
}

// One launch for one stage of several ideal-gas blocks of the same
// dimensionality: their StageArgs go to a device table (one slot per stage of
// the step, staged through pinned host memory; the host rewrites a slot only
// after the previous step's synchronisation), and each workgroup finds its
// block in a workgroup-to-entry map.  The k-chunk of every block is chosen so
// that the whole launch has >= 2048 workgroups.
// Copy a batched launch's argument table and workgroup map into the stage's
// slot of the staging buffers (grown, with the stream drained, when too small);
// a table or map equal to what the slot's device copy already holds is not
// copied again.
// A slot of set 0 skips the copy when its device copy already holds the
// table; the slots of a captured graph's set always record the copy (a capture
// runs nothing, so their device copies are only what the graph writes).
// Growing the buffers drops the captured graphs, whose nodes hold the old ones.
static int stage_batch_buffers(BatchBuffers& bb, int slot, const void* a, size_t na, const void* m, size_t nm,
                               char** ad, char** md)
{
    if (na > bb.args_slot || nm > bb.map_slot) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIP_TRY(hipStreamIsCapturing(g_state.stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(E4HIP_ERR_ARG, "batched argument tables outgrew their buffers during a graph capture");
        graph_invalidate();
        HIP_TRY(hipStreamSynchronize(g_state.stream));
        if (bb.args_dev) { HIP_TRY(hipFree(bb.args_dev)); HIP_TRY(hipHostFree(bb.args_host)); }
        if (bb.map_dev) { HIP_TRY(hipFree(bb.map_dev)); HIP_TRY(hipHostFree(bb.map_host)); }
        bb.args_slot = (std::max(na, bb.args_slot) * 2 + 255) / 256 * 256;
        bb.map_slot = (std::max(nm, bb.map_slot) * 2 + 255) / 256 * 256;
        HIP_TRY(hipMalloc(&bb.args_dev, bb.args_slot * BATCH_SLOTS));
        HIP_TRY(hipHostMalloc(&bb.args_host, bb.args_slot * BATCH_SLOTS));
        HIP_TRY(hipMalloc(&bb.map_dev, bb.map_slot * BATCH_SLOTS));
        HIP_TRY(hipHostMalloc(&bb.map_host, bb.map_slot * BATCH_SLOTS));
        for (int q = 0; q < BATCH_SLOTS; ++q) bb.args_valid[q] = bb.map_valid[q] = 0;
    }
    char* ah = (char*)bb.args_host + bb.args_slot * slot;
    char* mh = (char*)bb.map_host + bb.map_slot * slot;
    *ad = (char*)bb.args_dev + bb.args_slot * slot;
    *md = (char*)bb.map_dev + bb.map_slot * slot;
    const bool eager_set = slot < 6;
    if (!(eager_set && bb.args_valid[slot] == na && memcmp(ah, a, na) == 0)) {
        memcpy(ah, a, na);
        HIP_TRY(hipMemcpyAsync(*ad, ah, na, hipMemcpyHostToDevice, g_state.stream));
        if (eager_set) bb.args_valid[slot] = na;
    }
    if (!(eager_set && bb.map_valid[slot] == nm && memcmp(mh, m, nm) == 0)) {
        memcpy(mh, m, nm);
        HIP_TRY(hipMemcpyAsync(*md, mh, nm, hipMemcpyHostToDevice, g_state.stream));
        if (eager_set) bb.map_valid[slot] = nm;
    }
    return E4HIP_OK;
}

// The key of a batched launch for the autotuner: its blocks' shapes and the
// arithmetic.
std::vector<int> batch_tune_key(const std::vector<const BlockH*>& blks)
{
    std::vector<int> key = {-1};
    for (const BlockH* b : blks) {
        const std::vector<int> k = tune_key(*b);
        key.insert(key.end(), k.begin(), k.end());
    }
    return key;
}

// The number of k-chunks per block of a batched launch: the tuned one, else
// enough for >= 2048 workgroups in the launch.
long batch_chunks(const std::vector<const BlockH*>& blks)
{
    const std::vector<int> key = batch_tune_key(blks);
    for (const auto& t : g_state.kchunk_tuned)
        if (t.first == key) return t.second;
    long txy = 0;
    for (const BlockH* b : blks)
        txy += (long)((b->ni + TILE_X - 1) / TILE_X) * ((b->nj + TILE_Y - 1) / TILE_Y);
    return (2048 + txy - 1) / txy;
}

int launch_stage_batch(const std::vector<const BlockH*>& blks, const std::vector<unsigned int*>& counters,
                       int stage, double dt, int set, long force_chunks)
{
    const E4hipStepConfig& cfg = g_state.cfg;
    const int n = (int)blks.size();
    const long chunks = force_chunks > 0 ? force_chunks : batch_chunks(blks);
    std::vector<e4::StageArgs> args(n);
    std::vector<long> ntile(n);
    for (int e = 0; e < n; ++e) {
        const BlockH& b = *blks[e];
        int kc = 1;
        if (b.dims == 3) {
            kc = (int)((b.nk + chunks - 1) / chunks);
            if (kc < 4) kc = 4;
            if (kc > b.nk) kc = b.nk;
        }
        dim3 grid;
        int rc = stage_args_for(b, stage, dt, counters[e], kc, args[e], grid);
        if (rc) return rc;
        ntile[e] = (long)grid.x * grid.y * grid.z;
    }
    // the tile list, the largest blocks' tiles first (the reference's
    // localFluidBlocksBySize order, simcore.d:532-551); a persistent launch
    // takes them from the queue in this order
    std::vector<int> order(n);
    for (int e = 0; e < n; ++e) order[e] = e;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return blks[a]->ncells > blks[b]->ncells; });
    std::vector<int> map;
    for (int e : order) {
        args[e].wg0 = (int)map.size();
        map.insert(map.end(), ntile[e], e);
    }
    // the boundary-cell workgroups of all blocks, mapped after the stage's
    std::vector<int> bmap;
    for (int e = 0; e < n; ++e) {
        args[e].bwg0 = (int)bmap.size();
        const int nb = (args[e].fix_off[6] > 0) ? (args[e].fix_off[6] + 63) / 64 : 0;
        bmap.insert(bmap.end(), nb, e);
    }
    const size_t nstage_wg = map.size();
    map.insert(map.end(), bmap.begin(), bmap.end());
    char *ad = nullptr, *md = nullptr;
    int rc = stage_batch_buffers(g_state.batch, batch_slot(set, stage, blks[0]->dims), args.data(),
                                 sizeof(e4::StageArgs) * n, map.data(), sizeof(int) * map.size(), &ad, &md);
    if (rc) return rc;
    e4::BatchArgs B{};
    B.table = (unsigned long long)(uintptr_t)ad;
    B.wg_entry = (const int*)md;
    const bool def = args[0].high_order && args[0].apply_limiter && args[0].extrema_clipping;
    const int flux = kernel_flux(cfg.flux_calculator);
    hipEvent_t ev = nullptr;
    prof_begin(&ev);
    int nwg = (int)nstage_wg;
    // persistent: as many workgroups as the GPU holds at once (the stage
    // kernel's occupancy: three workgroups of four waves per compute unit in
    // the FP64 forms, four in the mixed form), taking tiles from the queue
    const int resident = g_state.cus * (cfg.precision_mode == E4HIP_PRECISION_MIXED ? 4 : 3);
    if (g_state.queue && nwg > resident) {
        B.queue = g_state.queue;
        B.ntiles = nwg;
        nwg = resident;
    }
    const int dims = blks[0]->dims;
    e4::BatchArgs Bb{};
    Bb.table = B.table;
    Bb.wg_entry = (const int*)md + nstage_wg;
    const int nbwg = (int)bmap.size();
    switch (cfg.precision_mode) {
    case E4HIP_PRECISION_FP64_EXACT:
        e4::launch_stage_batch_exact(B, nwg, def, g_state.stream, dims, flux);
        e4::launch_boundary_cells_batch_exact(Bb, nbwg, g_state.stream, dims, flux);
        break;
    case E4HIP_PRECISION_MIXED:
        e4::launch_stage_batch_mixed(B, nwg, def, g_state.stream, dims, flux);
        e4::launch_boundary_cells_batch_mixed(Bb, nbwg, g_state.stream, dims, flux);
        break;
    default:
        e4::launch_stage_batch_fast64(B, nwg, def, g_state.stream, dims, flux);
        e4::launch_boundary_cells_batch_fast64(Bb, nbwg, g_state.stream, dims, flux);
        break;
    }
    prof_end(0, ev);
    HIP_TRY(hipGetLastError());
    return E4HIP_OK;
}

// Stage sequencing of the thermally perfect path: the same U/dUdt buffer
// use as launch_stage; the primitive state goes P[0] -> P[1] -> P[0] ...
// Whether a mixture can use the fast modes' mixture form (tp_gas_device.h).
bool tp_mixture_form(const e4::TPGas& g)
{
    return g.shared != 0 && g.curve[0].nseg == e4::TP_FAST_NSEG;
}

// The thermally perfect kernels' arguments for one block's stage.
int tp_args_for(const BlockH& b, int stage, double dt, unsigned int* counters, e4::TPArgs& A)
{
    const E4hipStepConfig& cfg = g_state.cfg;
    const int sch = cfg.gasdynamic_update_scheme;
    const int ns = nstages(sch);
    memset(&A, 0, sizeof(A));
    A.qs = b.qs;
    for (int dd = 0; dd < 3; ++dd) { A.F[dd] = b.F[dd]; A.fqs[dd] = b.fqs[dd]; }
    A.ni = b.ni; A.nj = b.nj; A.nk = (b.dims == 3) ? b.nk : 1; A.dims = b.dims;
    A.stage = stage;
    A.high_order = cfg.interpolation_order > 1;
    A.apply_limiter = cfg.apply_limiter;
    A.extrema_clipping = cfg.extrema_clipping;
    A.counters = counters;
    A.gas = b.is_tp() ? g_state.gases[b.gm].tp_dev : nullptr;
    A.ig = g_state.gases[b.gm].ideal;
    A.guess = (cfg.learned_eos_guess && b.is_tp()) ? g_state.gases[b.gm].guess_dev : nullptr;
    A.bc = b.bc_dev;
    A.geo = b.geo;
    for (int dd = 0; dd < 3; ++dd) A.fgeo[dd] = b.fgeo[dd];
    for (int f = 0; f < 6; ++f) {
        if (b.bc_type[f] != E4HIP_BC_CONNECTED) continue;
        const BlockH* o = get_block(b.nbr_blk[f]);
        if (!o) return fail(E4HIP_ERR_HANDLE, "face %d is connected to a destroyed block", f);
        // the other block's primitive state at the start of this stage
        A.nbrP[f] = o->P[stage - 1];
        A.nbrGeo[f] = o->geo;
    }
    if (b.dims == 3) {
        A.vol_inv = 1.0 / (b.dx * b.dy * b.dz);
        A.area[0] = b.dy * b.dz; A.area[1] = b.dz * b.dx; A.area[2] = b.dx * b.dy;
    } else {
        A.vol_inv = 1.0 / (b.dx * b.dy);
        A.area[0] = b.dy; A.area[1] = b.dx; A.area[2] = 0.0;
    }
    A.w[0] = l2r2_weights(b.dx);
    A.w[1] = l2r2_weights(b.dy);
    A.w[2] = l2r2_weights(b.dz);
    double g1 = 1.0, g2 = 0.0, g3 = 0.0;
    stage_coefficients(sch, stage, g1, g2, g3);
    A.Unew = b.U[stage];
    if (stage == 1) {
        A.Uold = b.U[0];
        A.dout = (ns > 1) ? b.d[0] : nullptr;
    } else if (stage == 2) {
        A.d0 = b.d[0];
        A.Uold = (sch == E4HIP_UPDATE_DENMAN_RK3) ? b.U[1] : b.U[0];
        A.dout = (ns > 2) ? b.d[1] : nullptr;
    } else {
        A.d0 = b.d[0]; A.d1 = b.d[1];
        A.Uold = (sch == E4HIP_UPDATE_DENMAN_RK3) ? b.U[2] : b.U[0];
    }
    A.dt = dt; A.g1 = g1; A.g2 = g2; A.g3 = g3;
    A.c1 = dt * g1;
    A.Pin = b.P[stage - 1];
    A.Pout = b.P[stage];
    abort_fields(stage, A);
    // faces the boundary-face kernel recomputes: per slab, the face layers
    // (wall and the face after it; a simple-outflow-flux face)
    {
        int lo[3], hi[3], off[7];
        boundary_slabs(b, lo, hi, off, A.noghost, false);
        const int n[3] = {b.ni, b.nj, (b.dims == 3) ? b.nk : 1};
        int total = 0;
        for (int sl = 0; sl < 6; ++sl) {
            const int d = sl >> 1;
            const int L = (sl & 1) ? hi[d] : lo[d];   // cell layers: 2 wall, 1 outflow; face layers alike
            A.fix_layers[sl] = L;
            A.fix_off[sl] = total;
            total += L * n[(d + 1) % 3] * n[(d + 2) % 3];
        }
        A.fix_off[6] = total;
    }
    for (int f = 0; f < 6; ++f) { A.gx[f] = b.gx[f]; A.gS[f] = b.gS[f]; }
    A.axi = b.axi;
    A.src = b.src;
    const bool lean = b.is_tp() && !b.general && cfg.precision_mode == E4HIP_PRECISION_MIXED &&
                      tp_mixture_form(g_state.gases[b.gm].tp);
    A.retry = lean ? b.retry : nullptr;
    A.stats = g_state.stats;
    if (lean && !b.retry) return fail(E4HIP_ERR_ARG, "mixture block without its retry list");
    return shock_fields(b, A);
}

int launch_tp_stage(BlockH& b, int stage, double dt, unsigned int* counters)
{
    const E4hipStepConfig& cfg = g_state.cfg;
    e4::TPArgs A;
    int rc = tp_args_for(b, stage, dt, counters, A);
    if (rc) return rc;
    const int flux = kernel_flux(cfg.flux_calculator);
    hipEvent_t ev = nullptr;
    prof_begin(&ev);
    if (b.general) {
        const int nsp = b.is_tp() ? 5 : 0;
        const bool exact = cfg.precision_mode == E4HIP_PRECISION_FP64_EXACT ||
                           (b.is_tp() && !tp_mixture_form(g_state.gases[b.gm].tp));
        if (exact) e4::launch_gen_stage_exact(A, g_state.stream, flux, nsp);
        else if (cfg.precision_mode == E4HIP_PRECISION_MIXED) e4::launch_gen_stage_mixed(A, g_state.stream, flux, nsp);
        else e4::launch_gen_stage_fast64(A, g_state.stream, flux, nsp);
        prof_end(0, ev);
        HIP_TRY(hipGetLastError());
        return E4HIP_OK;
    }
    // the fast modes evaluate the mixture form, which needs three-segment curves
    // with shared break points; other mixtures run the reference-form kernels
    const bool mix_form = tp_mixture_form(g_state.gases[b.gm].tp);
    if (cfg.precision_mode == E4HIP_PRECISION_FP64_EXACT || !mix_form)
        e4::launch_tp_stage_exact(A, g_state.stream, flux);
    else if (cfg.precision_mode == E4HIP_PRECISION_MIXED) e4::launch_tp_stage_mixed(A, g_state.stream, flux);
    else e4::launch_tp_stage_fast64(A, g_state.stream, flux);
    prof_end(0, ev);
    HIP_TRY(hipGetLastError());
    return E4HIP_OK;
}

// Whether a mixture block's stage runs the reference-form (exact) kernels:
// in fp64_exact, or when its gas cannot use the fast modes' mixture form.
bool tp_exact_form(const BlockH& b)
{
    return g_state.cfg.precision_mode == E4HIP_PRECISION_FP64_EXACT || !tp_mixture_form(g_state.gases[b.gm].tp);
}

// One stage of several mixture blocks of one dimensionality (uniform cells,
// the same kernel form and learned-guess use): one launch per face direction
// over all blocks' faces, each block's boundary-face kernel where it has
// faces to recompute, one update launch over all blocks' cells.  The blocks'
// TPArgs go to a device table (staged like the ideal-gas batches), with a
// workgroup-to-block map per launch.
int launch_tp_stage_batch(const std::vector<BlockH*>& blks, const std::vector<unsigned int*>& counters, int stage,
                          double dt, int set)
{
    const E4hipStepConfig& cfg = g_state.cfg;
    const int n = (int)blks.size();
    const int dims = blks[0]->dims;
    std::vector<e4::TPArgs> args(n);
    for (int e = 0; e < n; ++e) {
        int rc = tp_args_for(*blks[e], stage, dt, counters[e], args[e]);
        if (rc) return rc;
    }
    std::vector<int> map;
    int wg[4] = {0, 0, 0, 0};
    size_t off[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        off[k] = map.size();
        if (k == 2 && dims != 3) continue;
        for (int e = 0; e < n; ++e) {
            const BlockH& b = *blks[e];
            const long nk = (dims == 3) ? b.nk : 1;
            long items;
            if (k == 0) items = (long)(b.ni + 1) * b.nj * nk;
            else if (k == 1) items = (long)b.ni * (b.nj + 1) * nk;
            else if (k == 2) items = (long)b.ni * b.nj * (b.nk + 1);
            else items = (long)b.ni * b.nj * nk;
            args[e].wg0[k] = (int)(map.size() - off[k]);
            map.insert(map.end(), (items + 127) / 128, e);
        }
        wg[k] = (int)(map.size() - off[k]);
    }
    char *ad = nullptr, *md = nullptr;
    int rc = stage_batch_buffers(g_state.tpbatch, batch_slot(set, stage, dims), args.data(),
                                 sizeof(e4::TPArgs) * n, map.data(), sizeof(int) * map.size(), &ad, &md);
    if (rc) return rc;
    e4::BatchArgs B[4];
    for (int k = 0; k < 4; ++k) {
        B[k] = e4::BatchArgs{};
        B[k].table = (unsigned long long)(uintptr_t)ad;
        B[k].wg_entry = (const int*)md + off[k];
    }
    std::vector<e4::TPArgs> per_block;
    for (int e = 0; e < n; ++e)
        if (args[e].fix_off[6] > 0) per_block.push_back(args[e]);
    const bool guess = args[0].guess != nullptr;
    const int flux = kernel_flux(cfg.flux_calculator);
    hipEvent_t ev = nullptr;
    prof_begin(&ev);
    if (tp_exact_form(*blks[0]))
        e4::launch_tp_stage_batch_exact(B, wg, dims, guess, per_block.data(), (int)per_block.size(), n, g_state.stream,
                                        flux);
    else if (cfg.precision_mode == E4HIP_PRECISION_MIXED)
        e4::launch_tp_stage_batch_mixed(B, wg, dims, guess, per_block.data(), (int)per_block.size(), n, g_state.stream,
                                        flux);
    else
        e4::launch_tp_stage_batch_fast64(B, wg, dims, guess, per_block.data(), (int)per_block.size(), n, g_state.stream,
                                         flux);
    prof_end(0, ev);
    HIP_TRY(hipGetLastError());
    return E4HIP_OK;
}

// Shock detection of one block for a stage (fluidblock.d:415-431): every
// block's indicators are swapped first (S[1] keeps the previous detection,
// which connected blocks' ghost cells carry), then each block detects from the
// stencil state of the stage into S[0].
int launch_shock_detect(const BlockH& b, int stage, double dt)
{
    const E4hipStepConfig& cfg = g_state.cfg;
    const bool exact = cfg.precision_mode == E4HIP_PRECISION_FP64_EXACT;
    if (b.general) {
        e4::TPArgs A;
        int rc = tp_args_for(b, stage, dt, nullptr, A);
        if (rc) return rc;
        if (exact || (b.is_tp() && !tp_mixture_form(g_state.gases[b.gm].tp))) e4::launch_gen_aux_exact(A, g_state.stream, 0, nullptr);
        else e4::launch_gen_aux_fast64(A, g_state.stream, 0, nullptr);
    } else if (b.is_tp()) {
        e4::TPArgs A;
        int rc = tp_args_for(b, stage, dt, nullptr, A);
        if (rc) return rc;
        if (exact || !tp_mixture_form(g_state.gases[b.gm].tp)) e4::launch_tp_shock_detect_exact(A, g_state.stream);
        else e4::launch_tp_shock_detect_fast64(A, g_state.stream);
    } else {
        e4::StageArgs A;
        dim3 grid;
        int rc = stage_args_for(b, stage, dt, nullptr, 0, A, grid);
        if (rc) return rc;
        if (exact) e4::launch_shock_detect_exact(A, b.dims, g_state.stream);
        else e4::launch_shock_detect_fast64(A, b.dims, g_state.stream);
    }
    HIP_TRY(hipGetLastError());
    return E4HIP_OK;
}

int alloc_block_buffers(BlockH& b)
{
    const size_t bytes = sizeof(double) * (size_t)b.qs * b.nq;
    const int ns = nstages(g_state.cfg.gasdynamic_update_scheme);
    if (b.has_P()) {
        const size_t pbytes = sizeof(double) * (size_t)b.qs * (8 + b.nsp);
        for (int t = 0; t <= ns; ++t) if (!b.P[t]) HIP_TRY(hipMalloc(&b.P[t], pbytes));
        const long nk = (b.dims == 3) ? b.nk : 1;
        const long nf[3] = {(long)(b.ni + 1) * b.nj * nk, (long)b.ni * (b.nj + 1) * nk,
                            (b.dims == 3) ? (long)b.ni * b.nj * (b.nk + 1) : 1};
        for (int dd = 0; dd < 3; ++dd) {
            b.fqs[dd] = (nf[dd] + 63) / 64 * 64;
            if (!b.F[dd]) HIP_TRY(hipMalloc(&b.F[dd], sizeof(double) * b.fqs[dd] * b.nq));
        }
        if (b.is_tp() && !b.general && !b.retry) {
            const size_t n = 4 + (size_t)nf[0] + nf[1] + ((b.dims == 3) ? nf[2] : 0);
            HIP_TRY(hipMalloc(&b.retry, sizeof(unsigned int) * n));
            HIP_TRY(hipMemset(b.retry, 0, sizeof(unsigned int) * 4));
        }
    }
    if (!b.bc_dev) HIP_TRY(hipMalloc(&b.bc_dev, sizeof(e4::BCTable)));
    for (int t = 0; t < 2; ++t) {
        if (b.S[t]) continue;
        HIP_TRY(hipMalloc(&b.S[t], (size_t)b.qs));
        HIP_TRY(hipMemset(b.S[t], 0, (size_t)b.qs));
    }
    for (int t = 0; t <= ns; ++t) if (!b.U[t]) HIP_TRY(hipMalloc(&b.U[t], bytes));
    for (int t = 0; t + 1 < ns; ++t) if (!b.d[t]) HIP_TRY(hipMalloc(&b.d[t], bytes));
    return E4HIP_OK;
}

int free_block_buffers(BlockH& b)
{
    for (auto& p : b.U) if (p) { HIP_TRY(hipFree(p)); p = nullptr; }
    for (auto& p : b.d) if (p) { HIP_TRY(hipFree(p)); p = nullptr; }
    for (auto& p : b.P) if (p) { HIP_TRY(hipFree(p)); p = nullptr; }
    for (auto& p : b.F) if (p) { HIP_TRY(hipFree(p)); p = nullptr; }
    if (b.retry) { HIP_TRY(hipFree(b.retry)); b.retry = nullptr; }
    for (auto& p : b.S) if (p) { HIP_TRY(hipFree(p)); p = nullptr; }
    for (auto& p : b.gx) if (p) { HIP_TRY(hipFree(p)); p = nullptr; }
    for (auto& p : b.gS) if (p) { HIP_TRY(hipFree(p)); p = nullptr; }
    if (b.axi) { HIP_TRY(hipFree(b.axi)); b.axi = nullptr; }
    if (b.src) { HIP_TRY(hipFree(b.src)); b.src = nullptr; }
    if (b.geo) { HIP_TRY(hipFree(b.geo)); b.geo = nullptr; }
    for (auto& p : b.fgeo) if (p) { HIP_TRY(hipFree(p)); p = nullptr; }
    if (b.bc_dev) { HIP_TRY(hipFree(b.bc_dev)); b.bc_dev = nullptr; }
    return E4HIP_OK;
}

// Copy the block's ghost-cell effects to its device table, after the work
// already queued on the library stream (which may read the old table).
int upload_bc(const BlockH& b)
{
    e4::BCTable t;
    memset(&t, 0, sizeof(t));
    for (int f = 0; f < 6; ++f) {
        t.type[f] = b.bc_type[f]; t.fs[f] = b.bc_fs[f]; t.nbr[f] = connection_of(b, f);
        memcpy(t.fs_tp[f], b.bc_fs_tp[f], sizeof(t.fs_tp[f]));
    }
    const GasModelH& g = g_state.gases[b.gm];
    if (g.type == GAS_IDEAL) {
        // gamma(Q) = Cp/Cv (gas_model.d:162) for the wall flux
        const double Cp = g.ideal.Rgas * g.ideal.gamma / (g.ideal.gamma - 1.0);
        t.wall_gamma = Cp / g.ideal.Cv;
    }
    // linear_extrap_prepare(len, len) (onedinterp.d:229-235)
    const double len[3] = {b.dx, b.dy, b.dz};
    for (int d = 0; d < 3; ++d) {
        t.ext_w[d][0] = (2.0 * len[d] + len[d]) / (len[d] + len[d]);
        t.ext_w[d][1] = -len[d] / (len[d] + len[d]);
    }
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    HIP_TRY(hipMemcpy(b.bc_dev, &t, sizeof(t), hipMemcpyHostToDevice));
    return E4HIP_OK;
}

} // namespace

extern "C" {

int e4hip_init(int device)
{
    if (g_state.init) return E4HIP_OK;
    HIP_TRY(hipSetDevice(device));
    if (device < 0 || device >= MAX_DEVICES) return fail(E4HIP_ERR_ARG, "device %d out of range", device);
    if (!scratch_dev[device]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_dev[device] = (unsigned long long*)p;
    }
    HIP_TRY(hipStreamCreateWithFlags(&g_state.stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&g_state.smax, sizeof(unsigned long long)));
    HIP_TRY(hipMalloc(&g_state.last_signal, sizeof(double)));
    HIP_TRY(hipMalloc(&g_state.queue, 2 * sizeof(unsigned int)));
    HIP_TRY(hipMemset(g_state.queue, 0, 2 * sizeof(unsigned int)));
    HIP_TRY(hipMalloc(&g_state.stats, sizeof(unsigned long long) * e4::STAT_LANES * e4::STAT_STRIDE));
    HIP_TRY(hipMemset(g_state.stats, 0, sizeof(unsigned long long) * e4::STAT_LANES * e4::STAT_STRIDE));
    {
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, device));
        g_state.cus = prop.multiProcessorCount;
    }
    g_state.device = device;
    g_state.init = true;
    int rc = ensure_counters(16);
    if (rc) return rc;
    return E4HIP_OK;
}

int e4hip_finalize(void)
{
    graph_invalidate();
    if (!g_state.init) return E4HIP_OK;
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    for (auto& b : g_state.blocks) if (b.alive) { free_block_buffers(b); b.alive = false; }
    g_state.blocks.clear();
    g_state.gases.clear();
    for (auto& pe : g_state.prof_pending) { hipEventDestroy(pe.a); hipEventDestroy(pe.b); }
    g_state.prof_pending.clear();
    if (g_state.counters_raw) HIP_TRY(hipFree(g_state.counters_raw));
    if (g_state.counters_host) HIP_TRY(hipHostFree(g_state.counters_host));
    g_state.counters_host = nullptr;
    g_state.counters_raw = g_state.abort = nullptr;
    for (BatchBuffers* bb : {&g_state.batch, &g_state.fill, &g_state.tpbatch}) {
        if (bb->args_dev) { HIP_TRY(hipFree(bb->args_dev)); HIP_TRY(hipHostFree(bb->args_host)); }
        if (bb->map_dev) { HIP_TRY(hipFree(bb->map_dev)); HIP_TRY(hipHostFree(bb->map_host)); }
        *bb = BatchBuffers();
    }
    if (g_state.smax) HIP_TRY(hipFree(g_state.smax));
    if (g_state.last_signal) HIP_TRY(hipFree(g_state.last_signal));
    if (g_state.queue) HIP_TRY(hipFree(g_state.queue));
    if (g_state.stats) HIP_TRY(hipFree(g_state.stats));
    g_state.counters = nullptr; g_state.smax = nullptr; g_state.last_signal = nullptr; g_state.queue = nullptr;
    g_state.counter_slots = 0;
    HIP_TRY(hipStreamDestroy(g_state.stream));
    g_state.stream = nullptr;
    g_state.init = false;
    return E4HIP_OK;
}

int e4hip_last_error(char* buf, int len)
{
    if (!buf || len <= 0) return E4HIP_ERR_ARG;
    snprintf(buf, (size_t)len, "%s", g_state.err.c_str());
    return E4HIP_OK;
}

void* e4hip_stream(void) { return (void*)g_state.stream; }

int e4hip_synchronize(void)
{
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    return E4HIP_OK;
}

int e4hip_gas_ideal(double mMass, double gamma, int* gm)
{
    graph_invalidate();
    if (!(mMass > 0.0) || !(gamma > 1.0) || !gm) return fail(E4HIP_ERR_ARG, "bad ideal-gas parameters");
    GasModelH g;
    g.mMass = mMass;
    g.ideal.Rgas = 8.31451 / mMass;                     // physical_constants.d:13, ideal_gas.d:65
    g.ideal.gamma = gamma;
    g.ideal.Cv = g.ideal.Rgas / (gamma - 1.0);
    g.ideal.Cvinv = 1.0 / g.ideal.Cv;
    g.ideal.gammaR = gamma * g.ideal.Rgas;
    g.ideal.CpCv = (g.ideal.Rgas * gamma / (gamma - 1.0)) / g.ideal.Cv;   // gas_model.d:162, ideal_gas.d:68
    g_state.gases.push_back(g);
    *gm = (int)g_state.gases.size() - 1;
    return E4HIP_OK;
}

// the sound speed of the constructed cell state of a mixture (fvcell.d:170:
// p = 100 kPa, T = 300 K, all of the first species; therm_perf_gas.d:345-355)
__global__ void tp_init_sound_speed_kernel(const e4::TPGas* g, double* a)
{
    e4::TPState<5> Q;
    Q.T = 300.0;
    for (int s = 0; s < 5; ++s) Q.Y[s] = (s == 0) ? 1.0 : 0.0;
    *a = e4::tp_sound_speed<e4::ModeExact, 5>(*g, Q);
}

int e4hip_gas_tp(int nsp, const double* M, int nseg, const double* Tbreaks, const double* Tblends,
                 const double* coeffs, int* gm)
{
    graph_invalidate();
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    if (nsp != 5) return fail(E4HIP_ERR_ARG, "the device path is built for 5 species (got %d)", nsp);
    if (nseg < 1 || nseg > e4::TP_NSEG_MAX || !M || !Tbreaks || !Tblends || !coeffs || !gm)
        return fail(E4HIP_ERR_ARG, "bad thermally perfect gas parameters");
    GasModelH g;
    g.type = GAS_TP;
    g.nsp = nsp;
    g.tp.nsp = nsp;
    for (int s = 0; s < nsp; ++s) {
        if (!(M[s] > 0.0)) return fail(E4HIP_ERR_ARG, "molecular mass must be positive");
        g.tp.R[s] = 8.31451 / M[s];                   // therm_perf_gas.d:68
        e4::CeaCurve& cv = g.tp.curve[s];
        cv.R = g.tp.R[s];
        cv.nseg = nseg;
        for (int q = 0; q <= nseg; ++q) cv.Tb[q] = Tbreaks[s * (nseg + 1) + q];
        for (int q = 0; q < nseg - 1; ++q) {
            cv.Tbl[q] = Tblends[s * (nseg - 1) + q];
            cv.rTbl[q] = (cv.Tbl[q] > 0.0) ? 1.0 / cv.Tbl[q] : 0.0;
        }
        for (int q = 0; q < nseg; ++q)
            for (int j = 0; j < 9; ++j) cv.c[q][j] = coeffs[(s * nseg + q) * 9 + j];
    }
    // the mixture form of the fast modes needs every species on the same breakpoints
    g.tp.shared = 1;
    for (int s = 1; s < nsp; ++s) {
        const e4::CeaCurve &a = g.tp.curve[0], &b = g.tp.curve[s];
        for (int q = 0; q <= nseg; ++q) if (a.Tb[q] != b.Tb[q]) g.tp.shared = 0;
        for (int q = 0; q < nseg - 1; ++q) if (a.Tbl[q] != b.Tbl[q]) g.tp.shared = 0;
    }
    for (int q = 0; q < nseg; ++q)
        for (int k = 0; k < 8; ++k)
            for (int s = 0; s < nsp; ++s) g.tp.RC[q][k][s] = g.tp.R[s] * g.tp.curve[s].c[q][k];
    HIP_TRY(hipMalloc(&g.tp_dev, sizeof(e4::TPGas)));
    HIP_TRY(hipMemcpy(g.tp_dev, &g.tp, sizeof(e4::TPGas), hipMemcpyHostToDevice));
    {
        double* da = nullptr;
        HIP_TRY(hipMalloc(&da, sizeof(double)));
        hipLaunchKernelGGL(tp_init_sound_speed_kernel, dim3(1), dim3(1), 0, g_state.stream, g.tp_dev, da);
        HIP_TRY(hipMemcpyAsync(&g.a_init, da, sizeof(double), hipMemcpyDeviceToHost, g_state.stream));
        HIP_TRY(hipStreamSynchronize(g_state.stream));
        HIP_TRY(hipFree(da));
    }
    g_state.gases.push_back(g);
    *gm = (int)g_state.gases.size() - 1;
    return E4HIP_OK;
}

int e4hip_gas_tp_set_guess_model(int gm, const float* params, const float* in_mean, const float* in_std,
                                 float T_mean, float T_std)
{
    graph_invalidate();
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    if (gm < 0 || gm >= (int)g_state.gases.size() || g_state.gases[gm].type != GAS_TP)
        return fail(E4HIP_ERR_HANDLE, "not a thermally perfect gas model: %d", gm);
    if (!params || !in_mean || !in_std) return fail(E4HIP_ERR_ARG, "null model arrays");
    GasModelH& g = g_state.gases[gm];
    // flat FP32 layout of the training kernels: W1[64][8], b1, W2[64][64], b2, W3[64][64], b3, w4[64], b4
    const float* W1 = params;
    const float* b1 = W1 + 64 * 8;
    const float* W2 = b1 + 64;
    const float* b2 = W2 + 64 * 64;
    const float* W3 = b2 + 64;
    const float* b3 = W3 + 64 * 64;
    const float* w4 = b3 + 64;
    const float* b4 = w4 + 64;
    std::vector<__bf16> w(64 * 32 + 2 * 64 * 64);
    for (int n = 0; n < 64; ++n)
        for (int k = 0; k < 32; ++k) w[n * 32 + k] = (__bf16)((k < 7) ? W1[n * 8 + k] : 0.0f);
    for (int q = 0; q < 64 * 64; ++q) { w[64 * 32 + q] = (__bf16)W2[q]; w[64 * 32 + 64 * 64 + q] = (__bf16)W3[q]; }
    if (!g.guess_w) HIP_TRY(hipMalloc(&g.guess_w, sizeof(__bf16) * w.size()));
    HIP_TRY(hipMemcpy(g.guess_w, w.data(), sizeof(__bf16) * w.size(), hipMemcpyHostToDevice));
    e4::GuessModel M;
    memset(&M, 0, sizeof(M));
    for (int f = 0; f < 7; ++f) { M.in_mean[f] = in_mean[f]; M.in_std[f] = in_std[f]; }
    M.T_mean = T_mean; M.T_std = T_std;
    for (int n = 0; n < 64; ++n) { M.b1[n] = b1[n]; M.b2[n] = b2[n]; M.b3[n] = b3[n]; M.w4[n] = w4[n]; }
    M.b4 = b4[0];
    M.W1 = g.guess_w;
    M.W2 = g.guess_w + 64 * 32;
    M.W3 = g.guess_w + 64 * 32 + 64 * 64;
    if (!g.guess_dev) HIP_TRY(hipMalloc(&g.guess_dev, sizeof(e4::GuessModel)));
    HIP_TRY(hipMemcpy(g.guess_dev, &M, sizeof(M), hipMemcpyHostToDevice));
    return E4HIP_OK;
}

int e4hip_tp_guess(int gm, int n, const double* x, double* T)
{
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    if (gm < 0 || gm >= (int)g_state.gases.size() || !g_state.gases[gm].guess_dev)
        return fail(E4HIP_ERR_HANDLE, "gas model %d has no guess model", gm);
    if (n <= 0 || !x || !T) return fail(E4HIP_ERR_ARG, "bad arguments");
    double *dx = nullptr, *dT = nullptr;
    HIP_TRY(hipMalloc(&dx, sizeof(double) * 7 * n));
    HIP_TRY(hipMalloc(&dT, sizeof(double) * n));
    HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * 7 * n, hipMemcpyHostToDevice, g_state.stream));
    e4::launch_mlp_guess(g_state.gases[gm].guess_dev, dx, n, dT, g_state.stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(T, dT, sizeof(double) * n, hipMemcpyDeviceToHost, g_state.stream));
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    HIP_TRY(hipFree(dx));
    HIP_TRY(hipFree(dT));
    return E4HIP_OK;
}

int e4hip_tp_decode_stats(long long* energy_evals, long long* fallbacks, int reset)
{
    return read_stats(e4::STAT_DECODE_EVALS, e4::STAT_DECODE_FALLBACKS, energy_evals, fallbacks, reset);
}

int e4hip_tp_face_stats(long long* energy_evals, long long* fallbacks, int reset)
{
    return read_stats(e4::STAT_FACE_EVALS, e4::STAT_FACE_FALLBACKS, energy_evals, fallbacks, reset);
}

} // extern "C"

namespace {

// Block creation: uniform cells of dx*dy*dz, or (geom != null) the geometry
// arrays of a block built from its grid (layout of e4hip_block_create_geometry).
int create_block(int gm, int dims, int ni, int nj, int nk, double dx, double dy, double dz, const double* geom,
                 int* blk)
{
    graph_invalidate();
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    if (gm < 0 || gm >= (int)g_state.gases.size()) return fail(E4HIP_ERR_HANDLE, "unknown gas model %d", gm);
    if (!blk || (dims != 2 && dims != 3)) return fail(E4HIP_ERR_ARG, "dims must be 2 or 3");
    if (dims == 2) nk = 1;
    if (ni < 2 || nj < 2 || (dims == 3 && nk < 2))
        return fail(E4HIP_ERR_ARG, "each active direction needs at least 2 cells");
    if (!(dx > 0) || !(dy > 0) || (dims == 3 && !(dz > 0))) return fail(E4HIP_ERR_ARG, "cell sizes must be positive");
    BlockH b;
    b.alive = true;
    b.gm = gm; b.dims = dims; b.ni = ni; b.nj = nj; b.nk = nk;
    b.dx = dx; b.dy = dy; b.dz = (dims == 3) ? dz : 1.0;
    b.ncells = (long)ni * nj * nk;
    b.qs = (b.ncells + 63) / 64 * 64;
    b.nsp = g_state.gases[gm].nsp;
    b.nq = (b.nsp > 1) ? 5 + b.nsp : 5;
    b.general = geom != nullptr;
    if (b.general) {
        // every cell needs a positive volume and positive lengths
        const long nc = b.ncells;
        for (long c = 0; c < nc; ++c)
            for (int q = 0; q <= dims; ++q)
                if (!(geom[q * nc + c] > 0.0))
                    return fail(E4HIP_ERR_ARG, "cell %ld has a non-positive %s", c, q ? "length" : "volume");
    }
    int rc = alloc_block_buffers(b);
    if (!rc) rc = upload_bc(b);
    if (!rc && b.general) {
        auto put = [&](double* dst, long stride, const double* src, int rows, long n) -> int {
            for (int q = 0; q < rows; ++q)
                HIP_TRY(hipMemcpy(dst + (size_t)q * stride, src + (size_t)q * n, sizeof(double) * n,
                                  hipMemcpyHostToDevice));
            return E4HIP_OK;
        };
        const long nc = b.ncells;
        const long nf[3] = {(long)(ni + 1) * nj * nk, (long)ni * (nj + 1) * nk, (long)ni * nj * (nk + 1)};
        rc = hipMalloc(&b.geo, sizeof(double) * 4 * b.qs) == hipSuccess ? E4HIP_OK
                                                                        : fail(E4HIP_ERR_HIP, "hipMalloc of the cell geometry failed");
        if (!rc) rc = put(b.geo, b.qs, geom, 4, nc);
        const double* fsrc = geom + 4 * nc;
        for (int d = 0; d < dims && !rc; ++d) {
            if (hipMalloc(&b.fgeo[d], sizeof(double) * 10 * b.fqs[d]) != hipSuccess)
                rc = fail(E4HIP_ERR_HIP, "hipMalloc of the face geometry failed");
            if (!rc) rc = put(b.fgeo[d], b.fqs[d], fsrc, 10, nf[d]);
            fsrc += 10 * nf[d];
        }
    }
    if (rc) { free_block_buffers(b); return rc; }
    g_state.blocks.push_back(b);
    *blk = (int)g_state.blocks.size() - 1;
    return E4HIP_OK;
}

} // namespace

extern "C" {

int e4hip_block_create(int gm, int dims, int ni, int nj, int nk, double dx, double dy, double dz, int* blk)
{
    return create_block(gm, dims, ni, nj, nk, dx, dy, dz, nullptr, blk);
}

int e4hip_block_create_geometry(int gm, int dims, int ni, int nj, int nk, const double* geom, int cartesian, int* blk)
{
    if (!geom) return fail(E4HIP_ERR_ARG, "null geometry");
    if (cartesian) return create_block(gm, dims, ni, nj, nk, geom[0], geom[1], (dims == 3) ? geom[2] : 1.0, nullptr, blk);
    return create_block(gm, dims, ni, nj, nk, 1.0, 1.0, 1.0, geom, blk);
}

int e4hip_block_destroy(int blk)
{
    graph_invalidate();
    BlockH* b = get_block(blk);
    if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blk);
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    int rc = free_block_buffers(*b);
    b->alive = false;
    return rc;
}

int e4hip_block_set_bc(int blk, int face, int type, const double* fs)
{
    graph_invalidate();
    BlockH* b = get_block(blk);
    if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blk);
    if (face < 0 || face > 5 || (b->dims == 2 && face > 3)) return fail(E4HIP_ERR_ARG, "bad face %d", face);
    if (type < E4HIP_BC_SLIP_WALL || type > E4HIP_BC_INFLOW_PROFILE || type == E4HIP_BC_CONNECTED)
        return fail(E4HIP_ERR_ARG, "bad bc type %d", type);
    const int dir = (face == E4HIP_FACE_EAST || face == E4HIP_FACE_WEST) ? 0
                  : (face == E4HIP_FACE_NORTH || face == E4HIP_FACE_SOUTH) ? 1 : 2;
    const long nfc = (long)((dir == 0) ? b->nj : b->ni) * ((dir == 2) ? b->nj : b->nk);
    const int np = 8 + b->nsp;
    auto ghost_buffer = [&](std::vector<double>& h) -> int {
        HIP_TRY(hipStreamSynchronize(g_state.stream));
        if (!b->gx[face]) HIP_TRY(hipMalloc(&b->gx[face], sizeof(double) * (size_t)np * 2 * nfc));
        HIP_TRY(hipMemcpy(b->gx[face], h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        return E4HIP_OK;
    };
    if (type == E4HIP_BC_INFLOW_PROFILE) {
        // fs: [2][face cells][9 + nsp] flow states (rho, p, T, u, a, vx, vy, vz, S,
        // Y...), layer-major, in-face index a + b * na
        if (!fs) return fail(E4HIP_ERR_ARG, "a static profile needs the ghost flow states");
        const int w = 9 + b->nsp;
        std::vector<double> h((size_t)np * 2 * nfc, 0.0);
        std::vector<unsigned char> S(2 * nfc, 0);
        static const int src_of[8] = {0, 5, 6, 7, 3, 2, 1, 4};   // P_* field <- fs column
        for (long g = 0; g < 2 * nfc; ++g) {
            const double* st = fs + g * w;
            if (!(st[0] > 0) || !(st[4] > 0) || !(b->is_tp() || st[3] > 0))
                return fail(E4HIP_ERR_ARG, "invalid profile state at ghost %ld", g);
            for (int q = 0; q < 8; ++q) h[(size_t)q * 2 * nfc + g] = st[src_of[q]];
            for (int s = 0; s < b->nsp; ++s) h[(size_t)(e4::P_Y + s) * 2 * nfc + g] = b->is_tp() ? st[9 + s] : 1.0;
            S[g] = st[8] != 0.0;
        }
        int rc = ghost_buffer(h);
        if (rc) return rc;
        if (!b->gS[face]) HIP_TRY(hipMalloc(&b->gS[face], 2 * nfc));
        HIP_TRY(hipMemcpy(b->gS[face], S.data(), 2 * nfc, hipMemcpyHostToDevice));
    }
    if (b->general && (type == E4HIP_BC_SLIP_WALL_1 || (type == E4HIP_BC_EXTRAPOLATE_1 && !b->is_tp())))
        return fail(E4HIP_ERR_ARG, "boundary kind %d is not available on a block built from its grid", type);
    if (type == E4HIP_BC_SLIP_WALL_1) {
        // one-sided reconstruction reaches three cells in from the wall
        const int n = (face == E4HIP_FACE_EAST || face == E4HIP_FACE_WEST) ? b->ni
                    : (face == E4HIP_FACE_NORTH || face == E4HIP_FACE_SOUTH) ? b->nj : b->nk;
        if (n < 3) return fail(E4HIP_ERR_ARG, "a wall without ghost cells needs at least 3 cells across it");
    }
    if (type == E4HIP_BC_INFLOW && b->is_tp()) {
        if (!fs) return fail(E4HIP_ERR_ARG, "inflow needs a flow state");
        if (!(fs[0] > 0) || !(fs[2] > 0) || !(fs[4] > 0)) return fail(E4HIP_ERR_ARG, "invalid inflow state");
        // fs = rho, p, T, u, a, vx, vy, vz, Y; stored in the primitive-array order
        double* t = b->bc_fs_tp[face];
        t[e4::P_RHO] = fs[0]; t[e4::P_P] = fs[1]; t[e4::P_T] = fs[2]; t[e4::P_U] = fs[3]; t[e4::P_A] = fs[4];
        t[e4::P_VX] = fs[5]; t[e4::P_VY] = fs[6]; t[e4::P_VZ] = fs[7];
        for (int q = 0; q < b->nsp; ++q) t[e4::P_Y + q] = fs[8 + q];
    } else if (type == E4HIP_BC_INFLOW) {
        if (!fs) return fail(E4HIP_ERR_ARG, "inflow needs a flow state");
        // fs = rho, p, T, u, a, vx, vy, vz
        Prim P;
        P.rho = fs[0]; P.u = fs[3]; P.a = fs[4]; P.v0 = fs[5]; P.v1 = fs[6]; P.v2 = fs[7];
        if (!(P.rho > 0) || !(P.u > 0) || !(P.a > 0)) return fail(E4HIP_ERR_ARG, "invalid inflow state");
        b->bc_fs[face] = P;
        // the same state in the primitive-array order (blocks built from their grid)
        double* t = b->bc_fs_tp[face];
        t[e4::P_RHO] = fs[0]; t[e4::P_P] = fs[1]; t[e4::P_T] = fs[2]; t[e4::P_U] = fs[3]; t[e4::P_A] = fs[4];
        t[e4::P_VX] = fs[5]; t[e4::P_VY] = fs[6]; t[e4::P_VZ] = fs[7];
    }
    if (type == E4HIP_BC_EXTRAPOLATE_1) {
        // the face's two ghost layers, holding the constructed cell state
        // (p = 100 kPa, T = 300 K, first species, at rest) until the first
        // fill; the fill does not write the sound speed
        std::vector<double> h((size_t)np * 2 * nfc, 0.0);
        const GasModelH& g = g_state.gases[b->gm];
        auto fill = [&](int field, double v) { std::fill(h.begin() + field * 2 * nfc, h.begin() + (field + 1) * 2 * nfc, v); };
        fill(e4::P_T, 300.0);
        fill(e4::P_P, 100.0e3);
        // the constructed cell's sound speed (ideal: sqrt(gamma R T) at 300 K,
        // ideal_gas.d:137-144)
        fill(e4::P_A, b->is_tp() ? g.a_init : sqrt(g.ideal.gammaR * 300.0));
        fill(e4::P_RHO, 100.0e3 / ((b->is_tp() ? g.tp.R[0] : g.ideal.Rgas) * 300.0));
        fill(e4::P_Y, 1.0);
        int rc = ghost_buffer(h);
        if (rc) return rc;
    }
    b->bc_type[face] = type;
    b->nbr_blk[face] = -1;
    return upload_bc(*b);
}

int e4hip_block_set_axisymmetric(int blk, double x0, double y0)
{
    graph_invalidate();
    BlockH* b = get_block(blk);
    if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blk);
    if (b->dims != 2) return fail(E4HIP_ERR_ARG, "axisymmetric geometry is for 2D blocks");
    if (b->general) return fail(E4HIP_ERR_ARG, "axisymmetric geometry is for uniform blocks");
    if (!(y0 >= 0.0) || !std::isfinite(x0)) return fail(E4HIP_ERR_ARG, "the block must lie at y >= 0");
    // xyplane_quad_cell_properties (properties.d:177-237; volume = area *
    // centroid y, fvcell.d:314-351) and update_2D_geometric_data
    // (fvinterface.d:265-297; area = length * mid-point y) of the cells and
    // faces of vertices (x0 + i dx, y0 + j dy); vertex orders sfluidblock.d:512-600
    const int ni = b->ni, nj = b->nj;
    auto X = [&](int i) { return x0 + i * b->dx; };
    auto Y = [&](int j) { return y0 + j * b->dy; };
    const long nc = (long)ni * nj;
    std::vector<double> h(2 * nc + (long)(ni + 1) * nj + (long)ni * (nj + 1));
    e4::AxiGeom g = e4::axi_geom(h.data(), ni, nj);
    double* vol = const_cast<double*>(g.vol);
    double* axy = const_cast<double*>(g.areaxy);
    double* ai = const_cast<double*>(g.ai);
    double* aj = const_cast<double*>(g.aj);
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const double xA = X(i + 1), yA = Y(j), xB = X(i + 1), yB = Y(j + 1);
            const double xC = X(i), yC = Y(j + 1), xD = X(i), yD = Y(j);
            const double area = 0.5 * ((xB + xA) * (yB - yA) + (xC + xB) * (yC - yB) +
                                       (xD + xC) * (yD - yC) + (xA + xD) * (yA - yD));
            const double yc = -1.0 / (area * 6.0) *
                              ((xB - xA) * (yA * yA + yA * yB + yB * yB) + (xC - xB) * (yB * yB + yB * yC + yC * yC) +
                               (xD - xC) * (yC * yC + yC * yD + yD * yD) + (xA - xD) * (yD * yD + yD * yA + yA * yA));
            vol[(long)j * ni + i] = area * yc;
            axy[(long)j * ni + i] = area;
        }
    auto face_area = [](double xA, double yA, double xB, double yB) {
        const double LAB = sqrt((xB - xA) * (xB - xA) + (yB - yA) * (yB - yA));
        const double Ybar = 0.5 * (yA + yB);
        return LAB * Ybar;
    };
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i <= ni; ++i) ai[(long)j * (ni + 1) + i] = face_area(X(i), Y(j), X(i), Y(j + 1));
    for (int j = 0; j <= nj; ++j)
        for (int i = 0; i < ni; ++i) aj[(long)j * ni + i] = face_area(X(i + 1), Y(j), X(i), Y(j));
    for (long c = 0; c < nc; ++c)
        if (!(vol[c] > 0.0)) return fail(E4HIP_ERR_ARG, "a cell has no volume (on the axis)");
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    if (!b->axi) HIP_TRY(hipMalloc(&b->axi, sizeof(double) * h.size()));
    HIP_TRY(hipMemcpy(b->axi, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    return E4HIP_OK;
}

int e4hip_block_set_source_terms(int blk, const double* Q)
{
    graph_invalidate();
    BlockH* b = get_block(blk);
    if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blk);
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    if (!Q) {
        if (b->src) { HIP_TRY(hipFree(b->src)); b->src = nullptr; }
        return E4HIP_OK;
    }
    for (long i = 0; i < (long)b->nq * b->ncells; ++i)
        if (!std::isfinite(Q[i])) return fail(E4HIP_ERR_ARG, "source term %ld is not finite", i);
    if (!b->src) {
        HIP_TRY(hipMalloc(&b->src, sizeof(double) * (size_t)b->nq * b->qs));
        HIP_TRY(hipMemset(b->src, 0, sizeof(double) * (size_t)b->nq * b->qs));
    }
    for (int q = 0; q < b->nq; ++q)
        HIP_TRY(hipMemcpy(b->src + (size_t)q * b->qs, Q + (size_t)q * b->ncells, sizeof(double) * b->ncells,
                          hipMemcpyHostToDevice));
    return E4HIP_OK;
}

int e4hip_block_connect(int blk, int face, int other, int other_face, int orientation)
{
    graph_invalidate();
    BlockH* b = get_block(blk);
    BlockH* o = get_block(other);
    if (!b || !o) return fail(E4HIP_ERR_HANDLE, "unknown block %d or %d", blk, other);
    const int nf = (b->dims == 3) ? 6 : 4;
    if (face < 0 || face >= nf || other_face < 0 || other_face >= nf) return fail(E4HIP_ERR_ARG, "bad face");
    if (b->dims != o->dims) return fail(E4HIP_ERR_ARG, "connected blocks must have the same dimensionality");
    if (b->gm != o->gm) return fail(E4HIP_ERR_ARG, "connected blocks must share the gas model");
    if (b->general != o->general)
        return fail(E4HIP_ERR_ARG, "a block built from its grid connects only to blocks built from their grids");
    if (orientation < 0 || orientation > 3 || (b->dims == 2 && orientation != 0))
        return fail(E4HIP_ERR_ARG, "bad orientation %d", orientation);
    // every ghost cell must map into the other block: check the corners of the face
    const int n[3] = {b->ni, b->nj, (b->dims == 3) ? b->nk : 1};
    const int on[3] = {o->ni, o->nj, (o->dims == 3) ? o->nk : 1};
    const int ax0 = (face == E4HIP_FACE_NORTH || face == E4HIP_FACE_SOUTH || face >= E4HIP_FACE_TOP) ? 0 : 1;
    const int ax1 = (face >= E4HIP_FACE_TOP) ? 1 : 2;
    const int na = n[ax0], nb = (b->dims == 3) ? n[ax1] : 1;
    // the face sizes must agree: the mapped face of the other block has as many cells
    const int oax0 = (other_face == E4HIP_FACE_NORTH || other_face == E4HIP_FACE_SOUTH || other_face >= E4HIP_FACE_TOP) ? 0 : 1;
    const int oax1 = (other_face >= E4HIP_FACE_TOP) ? 1 : 2;
    const long cells = (long)na * nb, ocells = (long)on[oax0] * ((b->dims == 3) ? on[oax1] : 1);
    if (cells != ocells) return fail(E4HIP_ERR_ARG, "connected faces have different cell counts");
    for (int layer = 1; layer <= 2; ++layer)
        for (int ca = 0; ca < 2; ++ca)
            for (int cb = 0; cb < 2; ++cb) {
                int src[3];
                e4::connection_source(b->dims, face, other_face, orientation, on, layer, ca ? na - 1 : 0,
                                      cb ? nb - 1 : 0, src);
                for (int d = 0; d < 3; ++d)
                    if (src[d] < 0 || src[d] >= on[d])
                        return fail(E4HIP_ERR_ARG, "face sizes do not match under orientation %d", orientation);
            }
    b->bc_type[face] = E4HIP_BC_CONNECTED;
    b->nbr_blk[face] = other;
    b->nbr_face[face] = other_face;
    b->nbr_orient[face] = orientation;
    return upload_bc(*b);
}

int e4hip_block_ncells(int blk, long* ncells, int* nq)
{
    BlockH* b = get_block(blk);
    if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blk);
    if (ncells) *ncells = b->ncells;
    if (nq) *nq = b->nq;
    return E4HIP_OK;
}

static int run_decode(const BlockH& b, double* fs_dev, unsigned int* nbad)
{
    hipLaunchKernelGGL(e4::reset_counters_kernel, dim3(1), dim3(64), 0, g_state.stream, g_state.counters,
                       e4::COUNTER_SLOT, nullptr);
    e4::DecodeArgs D;
    D.U = b.U[0]; D.fs = fs_dev; D.qs = b.qs; D.ncells = b.ncells;
    D.gas = g_state.gases[b.gm].ideal; D.counters = g_state.counters;
    hipLaunchKernelGGL(e4::decode_kernel, dim3((unsigned)((b.ncells + 255) / 256)), dim3(256), 0, g_state.stream, D);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(nbad, g_state.counters, sizeof(unsigned int), hipMemcpyDeviceToHost, g_state.stream));
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    return E4HIP_OK;
}

static e4::TPArgs tp_args_for(const BlockH& b)
{
    e4::TPArgs A;
    memset(&A, 0, sizeof(A));
    A.qs = b.qs; A.ni = b.ni; A.nj = b.nj; A.nk = (b.dims == 3) ? b.nk : 1; A.dims = b.dims;
    for (int d = 0; d < 3; ++d) A.fqs[d] = b.fqs[d];
    A.gas = g_state.gases[b.gm].tp_dev;
    A.counters = g_state.counters;
    A.stats = g_state.stats;
    return A;
}

int e4hip_block_upload(int blk, const double* U, const double* Tguess)
{
    graph_invalidate();
    BlockH* b = get_block(blk);
    if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blk);
    if (!U) return fail(E4HIP_ERR_ARG, "null U");
    if (b->is_tp() && !Tguess) return fail(E4HIP_ERR_ARG, "a thermally perfect gas needs the temperature guess");
    // an uploaded flow state carries no shock indicator (flowstate.d:44, S = 0)
    for (auto* sp : b->S) if (sp) HIP_TRY(hipMemsetAsync(sp, 0, (size_t)b->qs, g_state.stream));
    for (int q = 0; q < b->nq; ++q)
        HIP_TRY(hipMemcpyAsync(b->U[0] + (size_t)q * b->qs, U + (size_t)q * b->ncells, sizeof(double) * b->ncells,
                               hipMemcpyHostToDevice, g_state.stream));
    if (b->is_tp()) {
        double* Tg = nullptr;
        HIP_TRY(hipMalloc(&Tg, sizeof(double) * b->ncells));
        HIP_TRY(hipMemcpyAsync(Tg, Tguess, sizeof(double) * b->ncells, hipMemcpyHostToDevice, g_state.stream));
        hipLaunchKernelGGL(e4::reset_counters_kernel, dim3(1), dim3(64), 0, g_state.stream, g_state.counters,
                       e4::COUNTER_SLOT, nullptr);
        e4::TPArgs A = tp_args_for(*b);
        A.Pout = b->P[0];
        A.guess = g_state.cfg.learned_eos_guess ? g_state.gases[b->gm].guess_dev : nullptr;
        hipEvent_t ev = nullptr;
        prof_begin(&ev);
        if (g_state.cfg.precision_mode == E4HIP_PRECISION_FP64_EXACT || !tp_mixture_form(g_state.gases[b->gm].tp))
            e4::launch_tp_decode_exact(A, b->U[0], Tg, g_state.stream);
        else e4::launch_tp_decode_fast64(A, b->U[0], Tg, g_state.stream);
        prof_end(2, ev);
        HIP_TRY(hipGetLastError());
        unsigned int cnt[e4::COUNTER_SLOT] = {};
        HIP_TRY(hipMemcpyAsync(cnt, g_state.counters, sizeof(cnt), hipMemcpyDeviceToHost, g_state.stream));
        HIP_TRY(hipStreamSynchronize(g_state.stream));
        HIP_TRY(hipFree(Tg));
        unsigned int nbad = cnt[0];
        if (nbad) return fail(E4HIP_ERR_DECODE, "%u cells of block %d do not decode to a valid state", nbad, blk);
        return E4HIP_OK;
    }
    unsigned int nbad = 0;
    if (b->general) {
        hipLaunchKernelGGL(e4::reset_counters_kernel, dim3(1), dim3(64), 0, g_state.stream, g_state.counters,
                           e4::COUNTER_SLOT, nullptr);
        // the decode of an uploaded state is the reference's arithmetic in
        // every precision mode (as decode_kernel's for uniform blocks): the
        // stored flow state must not depend on the mode configured at upload
        e4::TPArgs A = tp_args_for(*b);
        A.Pout = b->P[0];
        A.ig = g_state.gases[b->gm].ideal;
        e4::launch_gen_aux_exact(A, g_state.stream, 1, b->U[0]);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&nbad, g_state.counters, sizeof(unsigned int), hipMemcpyDeviceToHost, g_state.stream));
        HIP_TRY(hipStreamSynchronize(g_state.stream));
    } else {
        int rc = run_decode(*b, nullptr, &nbad);
        if (rc) return rc;
    }
    if (nbad) return fail(E4HIP_ERR_DECODE, "%u cells of block %d do not decode to a valid state", nbad, blk);
    return E4HIP_OK;
}

int e4hip_block_download(int blk, double* U, double* fs)
{
    BlockH* b = get_block(blk);
    if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blk);
    if (U) {
        for (int q = 0; q < b->nq; ++q)
            HIP_TRY(hipMemcpyAsync(U + (size_t)q * b->ncells, b->U[0] + (size_t)q * b->qs, sizeof(double) * b->ncells,
                                   hipMemcpyDeviceToHost, g_state.stream));
    }
    if (fs && b->has_P()) {
        // primitive layout rho,vx,vy,vz,u,T,p,a,Y -> flow-state layout rho,p,T,u,a,vx,vy,vz,Y
        const int map[8] = {e4::P_RHO, e4::P_P, e4::P_T, e4::P_U, e4::P_A, e4::P_VX, e4::P_VY, e4::P_VZ};
        for (int q = 0; q < 8 + (b->is_tp() ? b->nsp : 0); ++q) {
            int src = (q < 8) ? map[q] : q;
            HIP_TRY(hipMemcpyAsync(fs + (size_t)q * b->ncells, b->P[0] + (size_t)src * b->qs,
                                   sizeof(double) * b->ncells, hipMemcpyDeviceToHost, g_state.stream));
        }
    } else if (fs) {
        double* fs_dev = nullptr;
        HIP_TRY(hipMalloc(&fs_dev, sizeof(double) * 8 * b->ncells));
        unsigned int nbad = 0;
        int rc = run_decode(*b, fs_dev, &nbad);
        if (rc) { hipFree(fs_dev); return rc; }
        HIP_TRY(hipMemcpyAsync(fs, fs_dev, sizeof(double) * 8 * b->ncells, hipMemcpyDeviceToHost, g_state.stream));
        HIP_TRY(hipStreamSynchronize(g_state.stream));
        HIP_TRY(hipFree(fs_dev));
    }
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    return E4HIP_OK;
}

int e4hip_config(const E4hipStepConfig* c)
{
    graph_invalidate();
    if (!c) return fail(E4HIP_ERR_ARG, "null config");
    // hlle is MHD only (fluxcalc.d:1197) and adaptive_hlle_roe takes it near shocks
    const int fc = c->flux_calculator;
    if (fc != E4HIP_FLUX_AUSMDV && fc != E4HIP_FLUX_HLLC && fc != E4HIP_FLUX_HANEL && fc != E4HIP_FLUX_EFM &&
        fc != E4HIP_FLUX_AUSM_PLUS_UP && fc != E4HIP_FLUX_ADAPTIVE_EFM_AUSMDV &&
        fc != E4HIP_FLUX_ADAPTIVE_HANEL_AUSMDV && fc != E4HIP_FLUX_ROE)
        return fail(E4HIP_ERR_ARG, "flux_calculator %d is not available on the device path", fc);
    if (!std::isfinite(c->shear_tolerance) || !std::isfinite(c->compression_tolerance))
        return fail(E4HIP_ERR_ARG, "shock tolerances must be finite");
    if (!std::isfinite(c->M_inf) || !(c->M_inf >= 0.0)) return fail(E4HIP_ERR_ARG, "M_inf must be finite and >= 0");
    if (c->gasdynamic_update_scheme < E4HIP_UPDATE_EULER || c->gasdynamic_update_scheme > E4HIP_UPDATE_DENMAN_RK3)
        return fail(E4HIP_ERR_ARG, "gasdynamic_update_scheme %d is not available", c->gasdynamic_update_scheme);
    if (c->interpolation_order != 1 && c->interpolation_order != 2)
        return fail(E4HIP_ERR_ARG, "interpolation_order must be 1 or 2");
    if (c->interpolate_in_local_frame != 1) return fail(E4HIP_ERR_ARG, "interpolate_in_local_frame must be true");
    if (c->thermo_interpolator != E4HIP_THERMO_RHOU) return fail(E4HIP_ERR_ARG, "thermo_interpolator must be rhou");
    if (c->precision_mode < E4HIP_PRECISION_FP64 || c->precision_mode > E4HIP_PRECISION_MIXED)
        return fail(E4HIP_ERR_ARG, "precision_mode %d not available", c->precision_mode);
    if (c->max_invalid_cells < 0) return fail(E4HIP_ERR_ARG, "max_invalid_cells must be >= 0");
    g_state.cfg = *c;
    for (auto& b : g_state.blocks)
        if (b.alive) { int rc = alloc_block_buffers(b); if (rc) return rc; }
    return E4HIP_OK;
}

// The buffers a step swaps between its stages, per block (PTRS_PER_BLOCK each).
static size_t step_pointers_to(const int* blks, int nblk, uintptr_t* v)
{
    static_assert(sizeof(BlockH::U) / sizeof(double*) + sizeof(BlockH::d) / sizeof(double*) +
                      sizeof(BlockH::P) / sizeof(double*) + sizeof(BlockH::S) / sizeof(unsigned char*) ==
                      PTRS_PER_BLOCK, "PTRS_PER_BLOCK");
    size_t i = 0;
    for (int n = 0; n < nblk; ++n) {
        const BlockH& b = *get_block(blks[n]);
        for (auto p : b.U) v[i++] = (uintptr_t)p;
        for (auto p : b.d) v[i++] = (uintptr_t)p;
        for (auto p : b.P) v[i++] = (uintptr_t)p;
        for (auto p : b.S) v[i++] = (uintptr_t)p;
    }
    return i;
}

static void set_step_pointers_from(const int* blks, int nblk, const uintptr_t* v)
{
    size_t i = 0;
    for (int n = 0; n < nblk; ++n) {
        BlockH& b = *get_block(blks[n]);
        for (auto& p : b.U) p = (double*)v[i++];
        for (auto& p : b.d) p = (double*)v[i++];
        for (auto& p : b.P) p = (double*)v[i++];
        for (auto& p : b.S) p = (unsigned char*)v[i++];
    }
}

static void step_pointers(const int* blks, int nblk, std::vector<uintptr_t>& v)
{
    const size_t o = v.size();
    v.resize(o + (size_t)PTRS_PER_BLOCK * nblk);
    step_pointers_to(blks, nblk, v.data() + o);
}

static void set_step_pointers(const int* blks, int nblk, const std::vector<uintptr_t>& v)
{
    set_step_pointers_from(blks, nblk, v.data());
}

// Everything one step puts on the library's stream (flag and counter reset,
// the stages' ghost fills, detector and stage kernels), with the host's
// swaps of the shock-indicator buffers.  counters: the block-stage counter
// slots; set: the batch staging set (0 eager, 1 and 2 captured graphs);
// batch: whether ideal-gas blocks may share a launch.
static int enqueue_step(const int* blks, int nblk, double dt, int ns, unsigned int* counters, int set, bool batch)
{
    int rc = 0;
    const int nslots = nblk * 3;
    hipLaunchKernelGGL(e4::reset_counters_kernel, dim3(1), dim3(1024), 0, g_state.stream, g_state.step_abort,
                       ABORT_WORDS + e4::COUNTER_SLOT * nslots, nullptr);
    for (int stage = 1; stage <= ns; ++stage) {
        // ghost cells that are stored rather than read through: the x_order-1
        // extrapolation (the other effects are applied on the fly); the ideal
        // blocks' faces in one launch when blocks are batched
        const bool exact = g_state.cfg.precision_mode == E4HIP_PRECISION_FP64_EXACT;
        std::vector<e4::StageArgs> fill_args;
        std::vector<int> fill_map;
        for (int n = 0; n < nblk; ++n) {
            BlockH& b = *get_block(blks[n]);
            int faces = 0;
            for (int f = 0; f < 6; ++f)
                if (b.gx[f] && b.bc_type[f] == E4HIP_BC_EXTRAPOLATE_1) faces |= 1 << f;
            if (!faces) continue;
            if (!b.has_P() && batch && g_state.cfg.batch_blocks) {
                e4::StageArgs A{};
                dim3 grid;
                rc = stage_args_for(b, stage, dt, nullptr, 0, A, grid);
                if (rc) return rc;
                const int e = (int)fill_args.size();
                fill_args.push_back(A);
                for (int f = 0; f < 6; ++f) {
                    if (!(faces & (1 << f))) continue;
                    const int dir = (f == E4HIP_FACE_EAST || f == E4HIP_FACE_WEST)
                                        ? 0 : ((f == E4HIP_FACE_NORTH || f == E4HIP_FACE_SOUTH) ? 1 : 2);
                    const long nfc = (long)((dir == 0) ? A.nj : A.ni) * ((dir == 2) ? A.nj : A.nk);
                    const int nw = (int)((nfc + 127) / 128);
                    for (int w = 0; w < nw; ++w) {
                        fill_map.push_back(e * 8 + f);
                        fill_map.push_back(w);
                    }
                }
                continue;
            }
            if (b.is_tp()) {
                e4::TPArgs A;
                rc = tp_args_for(b, stage, dt, nullptr, A);
                if (rc) return rc;
                if (exact || !tp_mixture_form(g_state.gases[b.gm].tp)) e4::launch_tp_extrap1_exact(A, faces, g_state.stream);
                else e4::launch_tp_extrap1_fast64(A, faces, g_state.stream);
            } else {
                e4::StageArgs A;
                dim3 grid;
                rc = stage_args_for(b, stage, dt, nullptr, 0, A, grid);
                if (rc) return rc;
                if (exact) e4::launch_extrapolate_ghosts_exact(A, faces, g_state.stream);
                else e4::launch_extrapolate_ghosts_fast64(A, faces, g_state.stream);
            }
            HIP_TRY(hipGetLastError());
        }
        if (!fill_args.empty()) {
            char *ad = nullptr, *md = nullptr;
            rc = stage_batch_buffers(g_state.fill, batch_slot(set, stage, 2), fill_args.data(),
                                     sizeof(e4::StageArgs) * fill_args.size(), fill_map.data(),
                                     sizeof(int) * fill_map.size(), &ad, &md);
            if (rc) return rc;
            e4::BatchArgs B{};
            B.table = (unsigned long long)(uintptr_t)ad;
            B.wg_entry = (const int*)md;
            const int nwg = (int)(fill_map.size() / 2);
            if (exact) e4::launch_extrapolate_ghosts_batch_exact(B, nwg, g_state.stream);
            else e4::launch_extrapolate_ghosts_batch_fast64(B, nwg, g_state.stream);
            HIP_TRY(hipGetLastError());
        }
        // adaptive_hanel_ausmdv turns the shock detector on (globalconfig.d:1942-1954);
        // it runs after the ghost-cell fill and before the fluxes (simcore.d:2352-2358)
        if (flux_detects_shocks(g_state.cfg.flux_calculator)) {
            for (int n = 0; n < nblk; ++n) {
                BlockH& b = *get_block(blks[n]);
                std::swap(b.S[0], b.S[1]);
            }
            for (int n = 0; n < nblk; ++n) {
                rc = launch_shock_detect(*get_block(blks[n]), stage, dt);
                if (rc) return rc;
            }
        }
        // ideal-gas blocks: one batched launch per dimensionality when there
        // are several; thermally perfect blocks: their kernel sequence each
        std::vector<const BlockH*> bl[2];
        std::vector<unsigned int*> bcount[2];
        std::vector<BlockH*> tl[2];
        std::vector<unsigned int*> tcount[2];
        for (int n = 0; n < nblk; ++n) {
            BlockH& b = *get_block(blks[n]);
            unsigned int* cnt = counters + e4::COUNTER_SLOT * (n * 3 + (stage - 1));
            if (b.has_P() && b.is_tp() && !b.general && batch && g_state.cfg.batch_blocks) {
                tl[b.dims - 2].push_back(&b);
                tcount[b.dims - 2].push_back(cnt);
            } else if (b.has_P()) {
                rc = launch_tp_stage(b, stage, dt, cnt);
                if (rc) return rc;
            } else if (batch) {
                bl[b.dims - 2].push_back(&b);
                bcount[b.dims - 2].push_back(cnt);
            } else {
                rc = launch_stage(b, stage, dt, cnt);
                if (rc) return rc;
            }
        }
        // mixture blocks: one launch per kernel over the blocks that share a
        // kernel form and learned-guess use, the others one by one
        for (int d = 0; d < 2; ++d) {
            std::vector<BlockH*> group;
            std::vector<unsigned int*> gcount;
            for (size_t e = 0; e < tl[d].size(); ++e) {
                const BlockH& a = *tl[d][0];
                const BlockH& b = *tl[d][e];
                const bool same = tp_exact_form(a) == tp_exact_form(b) &&
                                  (g_state.gases[a.gm].guess_dev != nullptr) == (g_state.gases[b.gm].guess_dev != nullptr);
                if (same) {
                    group.push_back(tl[d][e]);
                    gcount.push_back(tcount[d][e]);
                } else {
                    rc = launch_tp_stage(*tl[d][e], stage, dt, tcount[d][e]);
                    if (rc) return rc;
                }
            }
            if ((int)group.size() >= g_state.batch_min_blocks) {
                rc = launch_tp_stage_batch(group, gcount, stage, dt, set);
                if (rc) return rc;
            } else {
                for (size_t e = 0; e < group.size(); ++e) {
                    rc = launch_tp_stage(*group[e], stage, dt, gcount[e]);
                    if (rc) return rc;
                }
            }
        }
        for (int d = 0; d < 2; ++d) {
            if (g_state.cfg.batch_blocks && (int)bl[d].size() >= g_state.batch_min_blocks) {
                rc = launch_stage_batch(bl[d], bcount[d], stage, dt, set, 0);
                if (rc) return rc;
            } else {
                for (size_t e = 0; e < bl[d].size(); ++e) {
                    rc = launch_stage(*bl[d][e], stage, dt, bcount[d][e]);
                    if (rc) return rc;
                }
            }
        }
    }
    return E4HIP_OK;
}

// The outcome of a step from its abort flags and counters read back into h
// (ABORT_WORDS flags, then the block-stage slots): per-stage bad-cell counts
// summed over the blocks, and the status of simcore.d:2480-2493 -- a stage
// with a failed decode or a block over max_invalid_cells ends the step there.
static int step_outcome(const unsigned int* h, int nblk, int ns, int* n_bad_cells, bool* stopped)
{
    const int CS = e4::COUNTER_SLOT;
    const unsigned int* slots = h + ABORT_WORDS;
    int stage_bad[3] = {0, 0, 0};
    bool too_many = false, decode_fail = false;
    int first_blk = -1, first_stage = -1, first_count = 0;   // the first block over the limit
    g_state.last_bad.assign(3 * (size_t)nblk, 0);
    *stopped = false;
    for (int s = 0; s < ns && !*stopped; ++s) {
        for (int n = 0; n < nblk; ++n) {
            const unsigned int* c = &slots[CS * (n * 3 + s)];
            const unsigned int inv = c[0], dec = c[1];
            stage_bad[s] += (int)(inv + dec);
            g_state.last_bad[3 * (size_t)n + s] = (int)(inv + dec);
            if ((int)inv > g_state.cfg.max_invalid_cells) {
                if (!too_many) { first_blk = n; first_stage = s; first_count = (int)inv; }
                too_many = true;
            }
            if (dec) decode_fail = true;
        }
        if (too_many || decode_fail || h[ABORT_STRIDE * s]) *stopped = true;
    }
    if (n_bad_cells) for (int s = 0; s < 3; ++s) n_bad_cells[s] = stage_bad[s];
    if (decode_fail) return fail(E4HIP_ERR_DECODE, "Bad cell with negative mass.");
    if (too_many)
        return fail(E4HIP_ERR_BAD_CELLS, "Too many bad cells; go home. (stage %d: %d bad cells in block[%d])",
                    first_stage + 1, first_count, first_blk);
    if (*stopped) return fail(E4HIP_ERR_BAD_CELLS, "a stage stopped the step");
    return E4HIP_OK;
}

// After a step: one that ended without bad cells makes its last stage's state
// the current one (swap(U[0], U[stages]), simcore.d:2936-2941, and the same for
// a mixture's primitive state); one that stopped restores every buffer pointer
// the step started with, so U[0] holds the state before the step.
static void finish_step(const int* blks, int nblk, int ns, bool stopped, const uintptr_t* pre)
{
    if (stopped) {
        set_step_pointers_from(blks, nblk, pre);
        return;
    }
    for (int n = 0; n < nblk; ++n) {
        BlockH& b = *get_block(blks[n]);
        std::swap(b.U[0], b.U[ns]);
        if (b.has_P()) std::swap(b.P[0], b.P[ns]);
    }
}

// Capture of one step into a graph (batch staging set `set`); the graph is
// launched once here.
static int capture_step(const int* blks, int nblk, double dt, int ns, int set, const std::vector<uintptr_t>& key,
                        const std::vector<uintptr_t>& pre)
{
    HIP_TRY(hipStreamBeginCapture(g_state.stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(blks, nblk, dt, ns, g_state.counters, set, true);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g_state.stream, &graph);
    if (rc || e != hipSuccess) {
        if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
        set_step_pointers(blks, nblk, pre);
        if (rc) return rc;
        HIP_TRY(e);
    }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) set_step_pointers(blks, nblk, pre);
    HIP_TRY(e);
    StepGraph G;
    G.key = key;
    step_pointers(blks, nblk, G.post);
    G.exec = ex;
    G.set = set;
    G.last_use = g_state.step_counter;
    g_state.graphs.push_back(G);
    HIP_TRY(hipGraphLaunch(ex, g_state.stream));
    return E4HIP_OK;
}

int e4hip_explicit_increment(const int* blks, int nblk, double dt, int* n_bad_cells)
{
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    g_state.last_bad.clear();
    if (!blks || nblk <= 0) return fail(E4HIP_ERR_ARG, "no blocks");
    if (!(dt > 0.0)) return fail(E4HIP_ERR_ARG, "dt must be positive");
    for (int n = 0; n < nblk; ++n)
        if (!get_block(blks[n])) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blks[n]);
    const int ns = nstages(g_state.cfg.gasdynamic_update_scheme);
    int rc = ensure_counters(nblk * 3);
    if (rc) return rc;
    g_state.step_abort = g_state.abort;
    ++g_state.step_counter;
    const int nslots = nblk * 3;
    std::vector<uintptr_t> pre;
    step_pointers(blks, nblk, pre);
    // use_graphs: a step whose starting state (generation, blocks, dt, buffer
    // pointers) was that of one of the last two steps is captured into a graph
    // and launched; later steps from the same state replay it.  Consecutive
    // steps alternate between two states (the swap of U[0] and U[stages]), so
    // two graphs are kept, each with its own batch staging set.
    rc = autotune_blocks(blks, nblk, dt);
    if (rc) return rc;
    const bool graphs = g_state.cfg.use_graphs != 0 && !g_state.profiling;
    std::vector<uintptr_t> key;
    if (graphs) {
        uint64_t dtb;
        memcpy(&dtb, &dt, sizeof(dtb));
        key = {(uintptr_t)g_state.generation, (uintptr_t)nblk, (uintptr_t)dtb, (uintptr_t)g_state.counters_raw,
               (uintptr_t)g_state.batch.args_dev, (uintptr_t)g_state.batch.map_dev,
               (uintptr_t)g_state.fill.args_dev, (uintptr_t)g_state.fill.map_dev};
        for (int n = 0; n < nblk; ++n) key.push_back((uintptr_t)blks[n]);
        key.insert(key.end(), pre.begin(), pre.end());
    }
    StepGraph* hit = nullptr;
    if (graphs)
        for (auto& G : g_state.graphs)
            if (G.key == key) hit = &G;
    if (hit) {
        HIP_TRY(hipGraphLaunch(hit->exec, g_state.stream));
        hit->last_use = g_state.step_counter;
        set_step_pointers(blks, nblk, hit->post);
    } else {
        bool seen = false;
        for (auto& k : g_state.recent_keys) if (k == key) seen = true;
        if (graphs && seen) {
            // a free staging set, or the least recently used graph's
            int set = 1;
            if (g_state.graphs.size() >= 2) {
                size_t lru = (g_state.graphs[0].last_use <= g_state.graphs[1].last_use) ? 0 : 1;
                set = g_state.graphs[lru].set;
                (void)hipGraphExecDestroy(g_state.graphs[lru].exec);
                g_state.graphs.erase(g_state.graphs.begin() + lru);
            } else if (g_state.graphs.size() == 1) {
                set = 3 - g_state.graphs[0].set;
            }
            rc = capture_step(blks, nblk, dt, ns, set, key, pre);
            if (rc) return rc;
        } else {
            rc = enqueue_step(blks, nblk, dt, ns, g_state.counters, 0, true);
            if (rc) {
                set_step_pointers(blks, nblk, pre);
                return rc;
            }
            if (graphs) {
                g_state.recent_keys.push_back(key);
                if (g_state.recent_keys.size() > 2) g_state.recent_keys.erase(g_state.recent_keys.begin());
            }
        }
    }
    HIP_TRY(hipMemcpyAsync(g_state.counters_host, g_state.counters_raw,
                           sizeof(unsigned int) * (ABORT_WORDS + e4::COUNTER_SLOT * nslots), hipMemcpyDeviceToHost,
                           g_state.stream));
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    bool stopped = false;
    rc = step_outcome(g_state.counters_host, nblk, ns, n_bad_cells, &stopped);
    finish_step(blks, nblk, ns, stopped, pre.data());
    return rc;
}

} // extern "C"

namespace {

// The CFL-signal kernel of one block on the library stream: the block's
// largest signal as a bit pattern into *smax_bits (atomic max; the caller
// clears it first) and the last cell's signal into *last_signal.
int enqueue_cfl(const BlockH& b, unsigned long long* smax_bits, double* last_signal, bool profile = true)
{
    const int stringent = g_state.cfg.stringent_cfl != 0;
    long nwg = (b.ncells + 255) / 256;
    if (nwg > 4096) nwg = 4096;
    hipEvent_t ev = nullptr;
    if (profile) prof_begin(&ev);
    if (b.general) {
        e4::TPArgs A = ::tp_args_for(b);
        A.Pin = b.P[0];
        A.geo = b.geo;
        for (int d = 0; d < 3; ++d) A.fgeo[d] = b.fgeo[d];
        e4::launch_gen_cfl(A, stringent, smax_bits, last_signal, (int)nwg, g_state.stream);
    } else if (b.is_tp()) {
        hipLaunchKernelGGL(e4::tp_cfl_kernel, dim3((unsigned)nwg), dim3(256), 0, g_state.stream, b.P[0], b.qs, b.ni,
                           b.nj, b.nk, b.dims, stringent, b.dx, b.dy, b.dz, smax_bits, last_signal);
    } else {
        e4::CflArgs C;
        C.U = b.U[0]; C.qs = b.qs; C.ni = b.ni; C.nj = b.nj; C.nk = b.nk; C.dims = b.dims;
        C.stringent = stringent;
        C.iLength = b.dx; C.jLength = b.dy; C.kLength = b.dz;
        C.gas = g_state.gases[b.gm].ideal;
        C.smax_bits = smax_bits; C.last_signal = last_signal;
        hipLaunchKernelGGL(e4::cfl_kernel<e4::ModeExact>, dim3((unsigned)nwg), dim3(256), 0, g_state.stream, C);
    }
    if (profile) prof_end(1, ev);
    HIP_TRY(hipGetLastError());
    return E4HIP_OK;
}

// cfl_allow of the update scheme (fluidblock.d:798-803)
double cfl_allow_of_scheme()
{
    switch (nstages(g_state.cfg.gasdynamic_update_scheme)) {
    case 1: return 0.9;
    case 2: return 1.2;
    default: return 1.6;
    }
}

// A block's dt_allow and cfl_max from its largest signal and its last cell's
// signal (fluidblock.d:776-862): min over cells of cfl/signal == cfl/max(signal)
// and max of dt*signal == dt*max(signal), since rounding is monotone; the
// rescaling when check_cfl and cfl_max exceeds cfl_allow uses the last cell's
// signal, as fluidblock.d:856 does.
void block_dt(unsigned long long smax_bits, double last, double dt_current, double cfl_value, int check_cfl,
              double& dt_allow, double& cfl_max)
{
    const double cfl_allow = cfl_allow_of_scheme();
    double smax;
    memcpy(&smax, &smax_bits, sizeof(smax));
    dt_allow = cfl_value / smax;
    cfl_max = dt_current * smax;
    if (check_cfl && (cfl_max < 0.0 || cfl_max > cfl_allow)) {
        cfl_max = 0.5 * cfl_allow;
        dt_allow = cfl_max / last;
    }
}

} // namespace

extern "C" {

// Per-block, per-stage bad-cell counts of the last e4hip_explicit_increment
// (the counts the reference prints per block, simcore.d:2480-2487).
int e4hip_last_step_bad_cells(int nblk, int* counts)
{
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    if (!counts || nblk <= 0) return fail(E4HIP_ERR_ARG, "counts must hold nblk x 3 ints");
    if ((size_t)nblk * 3 != g_state.last_bad.size())
        return fail(E4HIP_ERR_ARG, "the last step had %d blocks, not %d (or did not run)",
                    (int)(g_state.last_bad.size() / 3), nblk);
    memcpy(counts, g_state.last_bad.data(), sizeof(int) * g_state.last_bad.size());
    return E4HIP_OK;
}

int e4hip_determine_dt(const int* blks, int nblk, double dt_current, double cfl_value, int check_cfl,
                       double* dt_allow_out, double* cfl_max_out)
{
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    if (!blks || nblk <= 0 || !dt_allow_out || !cfl_max_out) return fail(E4HIP_ERR_ARG, "bad arguments");
    double dt_allow_all = 1.7976931348623157e308, cfl_max_all = 0.0;
    for (int n = 0; n < nblk; ++n) {
        BlockH* b = get_block(blks[n]);
        if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blks[n]);
        hipLaunchKernelGGL(e4::reset_counters_kernel, dim3(1), dim3(64), 0, g_state.stream, g_state.counters, 0,
                           g_state.smax);
        int rc = enqueue_cfl(*b, g_state.smax, g_state.last_signal);
        if (rc) return rc;
        unsigned long long bits = 0;
        double last = 0.0;
        HIP_TRY(hipMemcpyAsync(&bits, g_state.smax, sizeof(bits), hipMemcpyDeviceToHost, g_state.stream));
        HIP_TRY(hipMemcpyAsync(&last, g_state.last_signal, sizeof(last), hipMemcpyDeviceToHost, g_state.stream));
        HIP_TRY(hipStreamSynchronize(g_state.stream));
        double dt_allow, cfl_max;
        block_dt(bits, last, dt_current, cfl_value, check_cfl, dt_allow, cfl_max);
        dt_allow_all = fmin(dt_allow_all, dt_allow);
        cfl_max_all = fmax(cfl_max_all, cfl_max);
    }
    *dt_allow_out = dt_allow_all;
    *cfl_max_out = cfl_max_all;
    return E4HIP_OK;
}

int e4hip_determine_dt_noalloc(const int* blks, int nblk, double dt_current, double cfl_value, int check_cfl,
                               double* dt_allow_out, double* cfl_max_out)
{
    if (!g_state.init) return fail(E4HIP_ERR_NOT_INIT, "e4hip_init not called");
    if (!blks || nblk <= 0 || !dt_allow_out || !cfl_max_out) return fail(E4HIP_ERR_ARG, "bad arguments");
    if (nblk > E4HIP_NOALLOC_MAX_BLOCKS)
        return fail(E4HIP_ERR_ARG, "e4hip_determine_dt_noalloc takes at most %d blocks (%d given)",
                    E4HIP_NOALLOC_MAX_BLOCKS, nblk);
    unsigned long long* sp = scratch_dev[g_state.device];
    if (!sp) return fail(E4HIP_ERR_NOT_INIT, "no scratch address for device %d", g_state.device);
    for (int n = 0; n < nblk; ++n)
        if (!get_block(blks[n])) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blks[n]);
    const int nw = 2 * nblk + 2;
    hipLaunchKernelGGL(noalloc_clear_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, g_state.stream, sp, nw);
    HIP_TRY(hipGetLastError());
    for (int n = 0; n < nblk; ++n) {
        int rc = enqueue_cfl(*get_block(blks[n]), sp + 2 * n, (double*)(sp + 2 * n + 1), false);
        if (rc) return rc;
    }
    double* res = (double*)(sp + 2 * nblk);
    hipLaunchKernelGGL(noalloc_dt_reduce_kernel, dim3(1), dim3(64), 0, g_state.stream, sp, nblk, dt_current,
                       cfl_value, check_cfl, cfl_allow_of_scheme(), res);
    HIP_TRY(hipGetLastError());
    double out[2];
    HIP_TRY(hipMemcpyAsync(out, res, sizeof(out), hipMemcpyDeviceToHost, g_state.stream));
    HIP_TRY(hipStreamSynchronize(g_state.stream));
    *dt_allow_out = out[0];
    *cfl_max_out = out[1];
    return E4HIP_OK;
}

int e4hip_profile_enable(int on)
{
    graph_invalidate();
    g_state.profiling = on != 0;
    return E4HIP_OK;
}

int e4hip_block_kchunk(int blk, int* kchunk)
{
    const BlockH* b = get_block(blk);
    if (!b) return fail(E4HIP_ERR_HANDLE, "unknown block %d", blk);
    if (!kchunk) return fail(E4HIP_ERR_ARG, "null kchunk");
    *kchunk = (b->dims == 3) ? kchunk_for(*b) : 1;
    return E4HIP_OK;
}

int e4hip_profile_read(int k, double* total_ms, long* launches)
{
    if (k < 0 || k > 2 || !total_ms || !launches) return fail(E4HIP_ERR_ARG, "bad arguments");
    if (!g_state.prof_pending.empty()) {
        HIP_TRY(hipStreamSynchronize(g_state.stream));
        for (auto& pe : g_state.prof_pending) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, pe.a, pe.b);
            g_state.prof_ms[pe.cls] += ms;
            g_state.prof_n[pe.cls] += 1;
            hipEventDestroy(pe.a);
            hipEventDestroy(pe.b);
        }
        g_state.prof_pending.clear();
    }
    *total_ms = g_state.prof_ms[k];
    *launches = g_state.prof_n[k];
    g_state.prof_ms[k] = 0.0;
    g_state.prof_n[k] = 0;
    return E4HIP_OK;
}

} // extern "C"
