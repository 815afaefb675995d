// stage_args.h -- launch parameters of the stage kernel, shared by the host
// API and the kernel translation units.
#pragma once
#include "gasdyn_device.h"
#include "tp_gas_device.h"
#include "learned_eos.h"

namespace e4 {

// 5: OutFlowBC_SimpleFlux, ghosts extrapolated (x_order 0) and the boundary
// face's flux from the interior cell; 6: WallBC_WithSlip1, no ghost-cell data,
// one-sided reconstruction and the wall flux; 7: OutFlowBC_SimpleExtrapolate
// with x_order 1.
// 8: InFlowBC_StaticProfile, a stored flow state per ghost cell.
enum { BC_SLIP_WALL = 0, BC_INFLOW = 1, BC_EXTRAPOLATE = 2, BC_PERIODIC = 3, BC_CONNECTED = 4,
       BC_OUTFLOW_SIMPLE_FLUX = 5, BC_SLIP_WALL_1 = 6, BC_EXTRAPOLATE_1 = 7, BC_INFLOW_PROFILE = 8 };
enum { FACE_NORTH = 0, FACE_EAST = 1, FACE_SOUTH = 2, FACE_WEST = 3, FACE_TOP = 4, FACE_BOTTOM = 5 };

// A face connected to a face of another block (full-face copy): the other
// block's cell counts, the face, the orientation, the dimensionality and the
// other block's quantity stride.
struct Connection {
    int n[3];
    int face, orient, dims;
    long qs;
};

// The cell of the other block that the ghost at (layer, a, b) of face f copies
// (full_face_copy.d:142-962).  In-face coordinates (a, b) are the index axes
// north/south (i, k), east/west (j, k), top/bottom (i, j); 2D faces use a.
// 2D: the coordinate runs reversed when f and g are both in {south, east} or
// both in {north, west}.  3D: with x~ = n - 1 - x, orientation o gives
// (a', b') = (a, b), (b~, a), (a~, b~), (b, a~) for o = 0..3, where o = 1 and 3
// trade places when f is right-handed (east, south, top: (a x b) along the
// outward normal), and the first output is reversed when f and g have the same
// handedness.  Layer L copies the cell at depth L-1 inside face g.
__host__ __device__ inline void connection_source(int dims, int f, int g, int o, const int* n, int layer, int a, int b,
                                                  int* src)
{
    // in-face axes and normal of face g, without indexed local arrays
    const bool ns_g = (g == FACE_NORTH || g == FACE_SOUTH), tb_g = (g == FACE_TOP || g == FACE_BOTTOM);
    const int A0 = (ns_g || tb_g) ? 0 : 1, A1 = tb_g ? 1 : 2, dn = ns_g ? 1 : (tb_g ? 2 : 0);
    const int nA0 = (A0 == 0) ? n[0] : n[1], nA1 = (A1 == 1) ? n[1] : n[2];
    const int nD = (dn == 0) ? n[0] : ((dn == 1) ? n[1] : n[2]);
    const bool plus_g = (g == FACE_NORTH || g == FACE_EAST || g == FACE_TOP);
    const bool right_f = (f == FACE_EAST || f == FACE_SOUTH || f == FACE_TOP);
    const bool right_g = (g == FACE_EAST || g == FACE_SOUTH || g == FACE_TOP);
    const int depth = plus_g ? nD - layer : layer - 1;
    int p, q = 0;
    if (dims == 2) {
        const bool se_f = (f == FACE_SOUTH || f == FACE_EAST), se_g = (g == FACE_SOUTH || g == FACE_EAST);
        p = (se_f == se_g) ? nA0 - 1 - a : a;
    } else {
        if (right_f && (o & 1)) o = (o + 2) & 3;
        switch (o) {
        case 0: p = a; q = b; break;
        case 1: p = nA0 - 1 - b; q = a; break;
        case 2: p = nA0 - 1 - a; q = nA1 - 1 - b; break;
        default: p = b; q = nA1 - 1 - a; break;
        }
        if (right_f == right_g) p = nA0 - 1 - p;
    }
    // (p, q) along (A0, A1), depth along dn; in 2D A1 is k and q = 0
    src[0] = (dn == 0) ? depth : p;
    src[1] = (dn == 1) ? depth : ((A0 == 1) ? p : q);
    src[2] = (dn == 2) ? depth : ((A1 == 2) ? q : 0);
}

// Ghost-cell effects of the six faces, kept in device memory so that the
// kernel reads them only where a stencil leaves the block.
struct BCTable {
    int type[6];
    int pad[2];
    Prim fs[6];                       // inflow ghost states (global frame)
    Connection nbr[6];                // connected faces
    double fs_tp[6][8 + TP_NSP_MAX];  // mixture inflow states in the primitive-array order (P_*)
    double wall_gamma;                // Cp/Cv of the ideal gas (gas_model.d:162) for the wall flux
    double ext_w[3][2];               // linear_extrap_prepare(len, len) weights per direction
};

struct StageArgs {
    const double* __restrict__ Us;    // state the fluxes are evaluated from (stencil)
    const double* __restrict__ Uold;  // base state of the update (cell-local)
    const double* __restrict__ d0;    // dUdt of stage 1 (stages 2, 3)
    const double* __restrict__ d1;    // dUdt of stage 2 (stage 3)
    double* __restrict__ dout;        // where this stage's dUdt is kept (or null)
    double* __restrict__ Unew;        // new state (may alias Uold)
    unsigned int* counters;           // [0] invalid cells, [1] decode failures
    long qs;                          // stride between conserved quantities
    int ni, nj, nk, kchunk;
    int stage;                        // 1, 2 or 3
    int high_order, apply_limiter, extrema_clipping;
    double c1;                        // dt*gamma_1 for stage 1
    double dt, g1, g2, g3;
    double vol_inv;
    double area[3];                   // i-, j-, k-face areas
    L2R2U wu[3];                      // l2r2 weights per direction (uniform grid)
    IdealGas gas;
    const BCTable* bc;
    const double* nbrU[6];            // stencil-state arrays of the blocks connected to each face
    int tiles_x, tiles_y;             // workgroup tiles along i and j
    int wg0;                          // first workgroup of this block in a batched launch
    int bwg0;                         // first workgroup of this block in a batched boundary-cell launch
    // shock indicators (adaptive_hanel_ausmdv, hanel): S of this stage's
    // detection, Sprev of the previous one (what the ghost cells were filled
    // with), the connected blocks' Sprev; Snew is written by the detector
    const unsigned char* S;
    const unsigned char* Sprev;
    const unsigned char* nbrS[6];
    unsigned char* Snew;
    double shear_tol, compression_tol;
    // flux calculator options: every face takes the near-shock calculator
    // (hanel, efm); that calculator is efm; FLUX_OTHER is roe (else
    // ausm_plus_up); ausm_plus_up's M_inf (globalconfig.d:591)
    int always_hanel, shock_efm, other_roe;
    double M_inf;
    // boundaries whose faces the interior code does not handle (one-sided
    // reconstruction, wall or outflow flux): the stage kernel leaves fix_lo[d]
    // and fix_hi[d] layers of cells at the two ends of direction d to the
    // boundary kernel, which covers fix_off[6] cells, slab by slab in the
    // order (d, side) = (0,-), (0,+), (1,-), (1,+), (2,-), (2,+) starting at
    // fix_off[slab]; noghost has bit f set when face f has no ghost-cell data
    int fix_lo[3], fix_hi[3];
    int fix_off[7];
    int noghost;
    // stored ghost cells of face f (static profile): [8][2 * face cells] in the
    // P_* field order, layer-major, in-face index a + b * na; gS their shock
    // indicators
    const double* gx[6];
    const unsigned char* gS[6];
    // 2D axisymmetric geometry (null: planar): cell volumes [ni*nj], cell
    // (x,y)-plane areas [ni*nj], i-face areas [(ni+1)*nj], j-face areas
    // [ni*(nj+1)], one array after the other
    const double* axi;
    // user-defined source terms per cell, constant in time, in U's layout
    // (null: none); added to the time derivative after the axisymmetric term
    const double* src;
    // bad-cell stop of a step (simcore.d:2480-2493): every kernel of a stage
    // whose abort_prev word is set (the stage before counted more invalid
    // cells in some block than max_invalid, or a decode failed) does nothing;
    // this stage's counts set abort_this
    const unsigned int* abort_prev;
    unsigned int* abort_this;
    unsigned int max_invalid;
};

// Whether the stage before this one stopped the step, read once per wave.
// Only the kernels of that stage write the word, and they have finished when
// a kernel of this stage starts (same stream), so a plain load sees it; the
// word sits alone in its cache line, away from the counters this stage's
// waves add to, so the load never queues behind their atomics.
__device__ __forceinline__ bool stage_aborted(const unsigned int* abort_prev)
{
    if (!abort_prev) return false;
    const unsigned int v = *abort_prev;
    return __builtin_amdgcn_readfirstlane(v) != 0u;
}

// The wave leader's share of a stage's counts: invalid cells and decode
// failures added to the block's counters; the step stops after this stage
// when the block's invalid count passes max_invalid (the atomic's old value
// tells which wave pushed it over) or any decode failed.
__device__ __forceinline__ void count_bad_cells(unsigned int* counters, unsigned int* abort_this, unsigned int max_invalid,
                                                unsigned int nbad, unsigned int nfail)
{
    if (nbad) {
        const unsigned int old = atomicAdd(&counters[0], nbad);
        if (abort_this && old + nbad > max_invalid) atomicOr(abort_this, 1u);
    }
    if (nfail) {
        atomicAdd(&counters[1], nfail);
        if (abort_this) atomicOr(abort_this, 1u);
    }
}

// Per-radian geometry of a 2D axisymmetric block, laid out as above.
struct AxiGeom {
    const double *vol, *areaxy, *ai, *aj;
};
__host__ __device__ inline AxiGeom axi_geom(const double* g, int ni, int nj)
{
    const long nc = (long)ni * nj;
    return AxiGeom{g, g + nc, g + 2 * nc, g + 2 * nc + (long)(ni + 1) * nj};
}

// A batched stage launch: one workgroup per (block, tile, k-chunk) over all
// blocks of the batch; workgroup w belongs to table entry wg_entry[w].  The
// table address is an integer so the struct is the same on host and device.
struct BatchArgs {
    unsigned long long table;         // const StageArgs[nentries] in device memory
    const int* wg_entry;              // entry of each tile (workgroup) of the launch
    unsigned int* queue;              // work queue (next tile, finished workgroups), or null
    int ntiles;                       // tiles of the launch (with a queue)
};

// Counters of one block and stage: [0] invalid cells, [1] decode failures,
// [2] decode energy evaluations, [3] decodes whose polish (from the learned
// guess or the stored temperature) fell back to the reference solve,
// [4] face energy evaluations, [5] face-state solves whose polish fell back
// (thermally perfect path).  A slot is 64 bytes: the slots that one stage's
// launches add to (the same stage of different blocks, three slots apart)
// never share a 128-byte cache line.  Words 2-5 are unused: the solver
// statistics go to the spread array below.
constexpr int COUNTER_SLOT = 16;

// The thermally perfect path's solver statistics, summed over the run:
// energy evaluations and fall-backs of the decodes and of the face-state
// solves.  Kept in STAT_LANES copies 128 bytes apart, each wave adding to the
// copy its workgroup index selects, so that the one atomic per wave of every
// kernel of a launch does not queue on one cache line (on a single line those
// atomics set the pace of the whole mixture path); the host adds the copies.
// Workgroups go to the GPU's eight XCDs (each with its own L2) round-robin in
// dispatch order.  xcd_contiguous maps dispatch index L of n to a logical
// index such that each XCD receives a contiguous run of logical workgroups:
// neighbouring tiles, which read each other's halo cells, then share one L2.
// A bijection on [0, n): the last n % 8 workgroups keep their index.
#ifndef E4_XCD_REMAP
#define E4_XCD_REMAP 1
#endif
constexpr int GPU_XCDS = 8;
__device__ __forceinline__ int xcd_contiguous(int L, int n)
{
    if (!E4_XCD_REMAP) return L;
    const int per = n / GPU_XCDS;
    return (L < per * GPU_XCDS) ? (L % GPU_XCDS) * per + L / GPU_XCDS : L;
}

constexpr int STAT_LANES = 64;
constexpr int STAT_STRIDE = 16;       // unsigned long longs: 128 bytes
enum { STAT_DECODE_EVALS = 0, STAT_DECODE_FALLBACKS = 1, STAT_FACE_EVALS = 2, STAT_FACE_FALLBACKS = 3 };

__device__ __forceinline__ void add_stat(unsigned long long* stats, int k, unsigned int v)
{
    if (stats && v) atomicAdd(&stats[(blockIdx.x % STAT_LANES) * STAT_STRIDE + k], (unsigned long long)v);
}

enum { P_RHO = 0, P_VX = 1, P_VY = 2, P_VZ = 3, P_U = 4, P_T = 5, P_P = 6, P_A = 7, P_Y = 8 };

struct TPArgs {
    const double* __restrict__ Pin;   // primitive state the stage starts from [8+nsp][qs]
    double* __restrict__ Pout;        // primitive state after the update
    const double* __restrict__ Uold;
    const double* __restrict__ d0;
    const double* __restrict__ d1;
    double* __restrict__ dout;
    double* __restrict__ Unew;
    double* F[3];                     // face fluxes [5+nsp][nfaces(dir)]
    unsigned int* counters;           // one slot of COUNTER_SLOT
    long qs;
    long fqs[3];
    int ni, nj, nk, dims;
    int stage;
    int high_order, apply_limiter, extrema_clipping;
    double c1, dt, g1, g2, g3;
    double vol_inv;
    double area[3];
    L2R2W w[3];
    const TPGas* gas;                 // device copy of the gas model (read-only)
    const GuessModel* guess;          // learned temperature guess of the decode, or null
    const BCTable* bc;                // boundary effects (device memory)
    const double* nbrP[6];            // primitive arrays of the connected blocks at the stage start
    // shock indicators, as in StageArgs
    const unsigned char* S;
    const unsigned char* Sprev;
    const unsigned char* nbrS[6];
    unsigned char* Snew;
    double shear_tol, compression_tol;
    int always_hanel, shock_efm, other_roe;   // flux calculator options, as in StageArgs
    double M_inf;
    // faces whose flux the face kernels do not compute as the reference does
    // (next to a wall without ghost data: the wall face and the one after it;
    // a simple-outflow-flux face), recomputed by the boundary-face kernel:
    // fix_off[6] faces, slab by slab in the order (d, side) = (0,-), (0,+),
    // (1,-), (1,+), (2,-), (2,+); fix_layers[slab] face layers each
    int fix_layers[6];
    // a batched launch over several mixture blocks: this block's first
    // workgroup in the launch of the face kernel of direction d (0-2) and of
    // the update kernel (3)
    int wg0[4];
    int fix_off[7];
    int noghost;                      // bit f: face f has no ghost-cell data
    // stored ghost cells of face f, [8+nsp][2 * face cells] (layer-major,
    // in-face index a + b * na): x_order-1 extrapolation, filled at the start
    // of a stage, or a static profile; gS the profile's shock indicators
    double* gx[6];
    const unsigned char* gS[6];
    const double* axi;                // 2D axisymmetric geometry, as in StageArgs
    const double* src;                // user-defined source terms, as in StageArgs
    const unsigned int* abort_prev;   // bad-cell stop of the step, as in StageArgs
    unsigned int* abort_this;
    unsigned int max_invalid;
    // a block built from its grid (null: uniform Cartesian cells): cell
    // geometry [4][qs] (volume, i/j/k lengths), face geometry of direction d
    // [10][fqs[d]] (area, n, t1, t2), the connected blocks' cell geometry
    const double* geo;
    const double* fgeo[3];
    const double* nbrGeo[6];
    IdealGas ig;                      // ideal-gas constants (ideal-gas blocks built from their grid)
    // mixed form: the faces whose short temperature solve did not settle,
    // handed by the face kernels to the retry kernel -- per direction d a
    // count at [d] and the face indices from [4 + first face slot of d]
    // (slots of direction 0, then 1, then 2); the update kernel clears the
    // counts.  Null in the other forms.
    unsigned int* retry;
    unsigned long long* stats;        // solver statistics (STAT_LANES copies, see add_stat)
};

// Launchers, one per arithmetic mode; each lives in its own translation unit
// (stage_exact.hip is compiled without FMA contraction, stage_fast.hip with it).
void launch_stage_exact(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux);
void launch_stage_fast64(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux);
void launch_stage_mixed(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux);
// batched: nwg workgroups over the blocks of one table (same dims; default
// reconstruction options or not, for all of them)
void launch_stage_batch_exact(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux);
// the x_order-1 ghost fill of several blocks' faces in one launch (128 face cells per workgroup)
void launch_extrapolate_ghosts_batch_exact(const BatchArgs& B, int nwg, hipStream_t stream);
void launch_extrapolate_ghosts_batch_fast64(const BatchArgs& B, int nwg, hipStream_t stream);
// the boundary-cell kernel over the blocks of a batched stage (64 cells per workgroup)
void launch_boundary_cells_batch_exact(const BatchArgs& B, int nwg, hipStream_t stream, int dims, int flux);
void launch_boundary_cells_batch_fast64(const BatchArgs& B, int nwg, hipStream_t stream, int dims, int flux);
void launch_boundary_cells_batch_mixed(const BatchArgs& B, int nwg, hipStream_t stream, int dims, int flux);
void launch_stage_batch_fast64(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux);
void launch_stage_batch_mixed(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux);
// thermally perfect mixture (five species): faces of every direction, then the
// update; decode of an uploaded state
void launch_tp_stage_exact(const TPArgs& A, hipStream_t stream, int flux);
void launch_tp_stage_batch_exact(const BatchArgs* B, const int* wg, int dims, bool guess, const TPArgs* per_block,
                                 int nper, int nblk, hipStream_t stream, int flux);
void launch_tp_stage_batch_fast64(const BatchArgs* B, const int* wg, int dims, bool guess, const TPArgs* per_block,
                                  int nper, int nblk, hipStream_t stream, int flux);
void launch_tp_stage_batch_mixed(const BatchArgs* B, const int* wg, int dims, bool guess, const TPArgs* per_block,
                                 int nper, int nblk, hipStream_t stream, int flux);
void launch_tp_stage_fast64(const TPArgs& A, hipStream_t stream, int flux);
void launch_tp_stage_mixed(const TPArgs& A, hipStream_t stream, int flux);
// shock indicators of a stage's stencil state (adaptive_hanel_ausmdv); the
// mixed mode uses the fast64 detector (FP64 arithmetic)
void launch_shock_detect_exact(const StageArgs& A, int dims, hipStream_t stream);
void launch_shock_detect_fast64(const StageArgs& A, int dims, hipStream_t stream);
void launch_tp_shock_detect_exact(const TPArgs& A, hipStream_t stream);
void launch_tp_shock_detect_fast64(const TPArgs& A, hipStream_t stream);
void launch_tp_decode_exact(const TPArgs& A, const double* U, const double* Tguess, hipStream_t stream);
// x_order-1 ghost cells of the mixture block's faces (face mask: bit f)
void launch_tp_extrap1_exact(const TPArgs& A, int faces, hipStream_t stream);
void launch_tp_extrap1_fast64(const TPArgs& A, int faces, hipStream_t stream);
void launch_tp_decode_fast64(const TPArgs& A, const double* U, const double* Tguess, hipStream_t stream);
// blocks built from their grid (nsp 0: ideal gas, 5: the mixture): one stage;
// what = 0 the shock detector, 1 the decode of an uploaded ideal-gas U into the
// flow-state array (U is rewritten where the low-temperature fallback re-encodes it)
void launch_gen_stage_exact(const TPArgs& A, hipStream_t stream, int flux, int nsp);
void launch_gen_stage_fast64(const TPArgs& A, hipStream_t stream, int flux, int nsp);
void launch_gen_stage_mixed(const TPArgs& A, hipStream_t stream, int flux, int nsp);
void launch_gen_aux_exact(const TPArgs& A, hipStream_t stream, int what, double* U);
void launch_gen_aux_fast64(const TPArgs& A, hipStream_t stream, int what, double* U);
void launch_gen_cfl(const TPArgs& A, int stringent, unsigned long long* smax_bits, double* last_signal, int nwg,
                    hipStream_t stream);

void launch_mlp_guess(const GuessModel* G, const double* x, int n, double* T, hipStream_t stream);
// the cells next to one-sided boundaries, after the stage kernel (A.fix_off[6] > 0)
void launch_boundary_cells_exact(const StageArgs& A, hipStream_t stream, int dims, int flux);
// x_order-1 ghost cells of an ideal-gas block's faces (face mask: bit f), at the start of a stage
void launch_extrapolate_ghosts_exact(const StageArgs& A, int faces, hipStream_t stream);
void launch_extrapolate_ghosts_fast64(const StageArgs& A, int faces, hipStream_t stream);
void launch_boundary_cells_fast64(const StageArgs& A, hipStream_t stream, int dims, int flux);
void launch_boundary_cells_mixed(const StageArgs& A, hipStream_t stream, int dims, int flux);

constexpr int TILE_X = 32;
constexpr int TILE_Y = 8;

} // namespace e4
