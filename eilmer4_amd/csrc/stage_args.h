// stage_args.h -- launch parameters of the stage kernel, shared by the host
// API and the kernel translation units.
#pragma once
#include "gasdyn_device.h"
#include "tp_gas_device.h"
#include "learned_eos.h"

namespace e4 {

enum { BC_SLIP_WALL = 0, BC_INFLOW = 1, BC_EXTRAPOLATE = 2, BC_PERIODIC = 3 };
enum { FACE_NORTH = 0, FACE_EAST = 1, FACE_SOUTH = 2, FACE_WEST = 3, FACE_TOP = 4, FACE_BOTTOM = 5 };

// Ghost-cell effects of the six faces, kept in device memory so that the
// kernel reads them only where a stencil leaves the block.
struct BCTable {
    int type[6];
    int pad[2];
    Prim fs[6];                       // inflow ghost states (global frame)
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
};

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
    unsigned int* counters;           // [0] invalid, [1] decode failures, [2] decode energy evaluations,
                                      // [3] cells whose learned guess fell back to the reference solve
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
    int bc_type[6];
    double bc_fs[6][8 + TP_NSP_MAX];  // ghost flow state of inflow faces: rho,p,T,u,a,vx,vy,vz,Y
};

// Launchers, one per arithmetic mode; each lives in its own translation unit
// (stage_exact.hip is compiled without FMA contraction, stage_fast.hip with it).
void launch_stage_exact(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux);
void launch_stage_fast64(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux);
void launch_stage_mixed(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux);
// thermally perfect mixture (five species): faces of every direction, then the
// update; decode of an uploaded state
void launch_tp_stage_exact(const TPArgs& A, hipStream_t stream, int flux);
void launch_tp_stage_fast64(const TPArgs& A, hipStream_t stream, int flux);
void launch_tp_decode_exact(const TPArgs& A, const double* U, const double* Tguess, hipStream_t stream);
void launch_tp_decode_fast64(const TPArgs& A, const double* U, const double* Tguess, hipStream_t stream);

void launch_mlp_guess(const GuessModel* G, const double* x, int n, double* T, hipStream_t stream);

constexpr int TILE_X = 32;
constexpr int TILE_Y = 8;

} // namespace e4
