/* e4hip.h -- C ABI of the MI355X explicit gas-dynamic update for structured
 * fluid blocks.  All entry points are extern "C", take plain pointers and
 * sizes, return an int status (0 = E4HIP_OK, < 0 = error; text via
 * e4hip_last_error) and never throw across the boundary.
 *
 * This is the drop-in for Eilmer4's
 *   gasdynamic_explicit_increment_with_fixed_grid()  (eilmer/simcore.d:2303-2952)
 * on one GPU, together with the pieces of its callers that touch block data:
 *   determine_time_step_size()                       (eilmer/simcore.d:1347-1452,
 *                                                      eilmer/fluidblock.d:776-862)
 *   block upload after init_simulation's decode      (eilmer/simcore.d:313-331)
 *   block download before write_solution_files       (eilmer/simcore.d:637-684)
 * The FFI precedent it follows is the reference's own GPU chemistry binding
 * (eilmer/cuda_gpu_chem.d:23-27: extern(C) launcher, plain double* arrays) and
 * the handle/int-status convention of gas/gas_cwrap.d:35-140.
 *
 * Array layouts: a block has ni*nj*nk active cells (nk = 1 in 2D), index
 * c = (k*nj + j)*ni + i (the reference's ijk order without the ghost layers,
 * eilmer/sfluidblock.d:205-209).  Conserved arrays are structure-of-arrays
 * [nq][ncells] with q = mass, x-momentum, y-momentum, z-momentum, total
 * energy, then rho*Y_s for each species of a multi-species gas
 * (eilmer/conservedquantities.d:18-50).  Flow-state arrays are
 * [8 + nsp][ncells]: rho, p, T, u, a, vel.x, vel.y, vel.z, massf[...]
 * (eilmer/flowstate.d:30-44, gas/gas_state.d:18-47).
 */
#ifndef E4HIP_H
#define E4HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes */
#define E4HIP_OK               0
#define E4HIP_ERR_ARG         -1  /* bad argument or unsupported combination */
#define E4HIP_ERR_HIP         -2  /* a HIP runtime call failed */
#define E4HIP_ERR_HANDLE      -3  /* unknown gas-model or block handle */
#define E4HIP_ERR_BAD_CELLS   -4  /* invalid cells above max_invalid_cells
                                     ("Too many bad cells", simcore.d:2480-2493) */
#define E4HIP_ERR_DECODE      -5  /* decode_conserved would throw: non-positive
                                     density or failed thermo (fvcell.d:570-589, 705-719) */
#define E4HIP_ERR_NOT_INIT    -6  /* e4hip_init not called */

/* ---- FluxCalculator (eilmer/fvcore.d:184-194; values kept, hllc appended) */
#define E4HIP_FLUX_AUSMDV      0
#define E4HIP_FLUX_HANEL       1   /* fluxcalc.d:532-623 */
#define E4HIP_FLUX_EFM         2   /* fluxcalc.d:626-820 (efmflx with exxef) */
#define E4HIP_FLUX_AUSM_PLUS_UP 3  /* fluxcalc.d:860-1037, with E4hipStepConfig.M_inf */
#define E4HIP_FLUX_ADAPTIVE_EFM_AUSMDV 4  /* EFM at shock points with small shear, AUSMDV
                                             elsewhere (fluxcalc.d:802-820); "adaptive" */
#define E4HIP_FLUX_ADAPTIVE_HANEL_AUSMDV 5  /* the reference's default (globalconfig.d:533):
                                              Hanel at shock points with small shear,
                                              AUSMDV elsewhere (fluxcalc.d:822-836) */
#define E4HIP_FLUX_ROE         8   /* fluxcalc.d:1202-1371 (Roe-Pike, Kermani-Plett fix) */
#define E4HIP_FLUX_HLLC        9   /* (no counterpart in the reference; Toro 2009 10.4) */
/* adaptive_hlle_roe (6) and hlle (7, MHD only, fluxcalc.d:1197) are refused. */

/* ---- GasdynamicUpdate (eilmer/fvcore.d:30-41) */
#define E4HIP_UPDATE_EULER       0
#define E4HIP_UPDATE_PC          1
#define E4HIP_UPDATE_MIDPOINT    2
#define E4HIP_UPDATE_CLASSIC_RK3 3
#define E4HIP_UPDATE_TVD_RK3     4
#define E4HIP_UPDATE_DENMAN_RK3  5

/* ---- InterpolateOption (eilmer/fvcore.d:155); the device path implements rhou,
 *      the reference default (globalconfig.d:565) */
#define E4HIP_THERMO_RHOU      1

/* ---- ghost-cell boundary effects (eilmer/bc/ghost_cell_effect/) */
#define E4HIP_BC_SLIP_WALL     0  /* WallBC_WithSlip: internal_copy_then_reflect.d:94-310 */
#define E4HIP_BC_INFLOW        1  /* InFlowBC_Supersonic: flow_state_copy.d:66-169 */
#define E4HIP_BC_EXTRAPOLATE   2  /* OutFlowBC_SimpleExtrapolate x_order 0: extrapolate_copy.d:82- */
#define E4HIP_BC_PERIODIC      3  /* ExchangeBC_FullFace of a block with itself across
                                     opposite faces: full_face_copy.d:1277-1357 */
#define E4HIP_BC_CONNECTED     4  /* ExchangeBC_FullFace to another block (set by
                                     e4hip_block_connect): full_face_copy.d:142-962, 1277-1357 */
#define E4HIP_BC_OUTFLOW_SIMPLE_FLUX 5  /* OutFlowBC_SimpleFlux (alias OutFlowBC_Simple, bc.lua:1124-1148):
                                     ghosts by x_order-0 extrapolation, the boundary face's flux
                                     from the interior cell (boundary_flux_effect.d:367-546) */
#define E4HIP_BC_SLIP_WALL_1   6  /* WallBC_WithSlip1 (bc.lua:554-577): no ghost-cell data; one-sided
                                     reconstruction onedinterp.d:165-1037 and the wall flux
                                     fluxcalc.d:130-298 (sfluidblock.d:2189-2345); needs >= 3 cells
                                     across the wall */
#define E4HIP_BC_EXTRAPOLATE_1 7  /* OutFlowBC_SimpleExtrapolate{xOrder=1}: extrapolate_copy.d:102-186 */
#define E4HIP_BC_INFLOW_PROFILE 8 /* InFlowBC_StaticProfile (bc.lua:996-1019): each ghost cell copies
                                     the flow state of its nearest profile point
                                     (flow_state_copy_from_profile.d, flowstate.d:589-736);
                                     the caller does the matching and passes the states */

/* ---- face names (eilmer/fvcore.d Face enum order) */
#define E4HIP_FACE_NORTH  0
#define E4HIP_FACE_EAST   1
#define E4HIP_FACE_SOUTH  2
#define E4HIP_FACE_WEST   3
#define E4HIP_FACE_TOP    4
#define E4HIP_FACE_BOTTOM 5

/* ---- arithmetic of the flux path */
#define E4HIP_PRECISION_FP64        0  /* plain form: FP64 throughout, FMA contraction and
                                          reciprocal-based division (~1 ulp per operation) */
#define E4HIP_PRECISION_FP64_EXACT  1  /* FP64 with the reference's roundings: IEEE division,
                                          no contraction, reference operation order */
#define E4HIP_PRECISION_MIXED       2  /* FP32 reconstruction and flux; FP64 residual
                                          accumulation, update and decode */

/* Step configuration: the LocalConfig fields the path reads
 * (eilmer/globalconfig.d:440-569, 875, 1008). */
typedef struct E4hipStepConfig {
    int flux_calculator;             /* E4HIP_FLUX_*; the reference's default is
                                        E4HIP_FLUX_ADAPTIVE_HANEL_AUSMDV (globalconfig.d:533) */
    int gasdynamic_update_scheme;    /* E4HIP_UPDATE_*; default pc */
    int interpolation_order;         /* 1 or 2; default 2 */
    int apply_limiter;               /* van Albada on/off; default 1 */
    int extrema_clipping;            /* default 1 */
    int interpolate_in_local_frame;  /* must be 1 (default) */
    int thermo_interpolator;         /* must be E4HIP_THERMO_RHOU (default) */
    int precision_mode;              /* E4HIP_PRECISION_* */
    int max_invalid_cells;           /* default 0 */
    int learned_eos_guess;           /* thermally perfect gas: start the decode's temperature
                                        solve from the learned guess (e4hip_gas_tp_set_guess_model)
                                        with an FP64 Newton polish; 0 = the reference's solve
                                        from the stored temperature */
    int batch_blocks;                /* 1: the ideal-gas blocks of one stage (of one
                                        dimensionality) go into one kernel launch when there
                                        are two or more; 0: one launch per block */
    double shear_tolerance;          /* adaptive_hanel_ausmdv: default 0.20 (globalconfig.d:585) */
    double compression_tolerance;    /* shock detector: default -0.30 (globalconfig.d:600) */
    int stringent_cfl;               /* CFL signal from the fastest normal velocity over the
                                        shortest cell length (fvcell.d:1642-1650); default 0
                                        (globalconfig.d:729) */
    int use_graphs;                  /* 1: a step that starts from the same state (blocks,
                                        boundaries, dt, buffers) as the step before is captured
                                        into a HIP graph, and later such steps replay it (fewer
                                        launches for many small blocks); any call that changes
                                        blocks, boundaries or the config drops the graph.
                                        Default 0 */
    double M_inf;                    /* ausm_plus_up's floor on the reference Mach number
                                        squared: default 0.01 (globalconfig.d:591) */
    int autotune;                    /* 1: before the first step of a block shape (and of a
                                        batched block set) under a new arithmetic, time the
                                        stage kernel at candidate k-chunk heights and keep the
                                        fastest (results do not depend on it).  Default 1 */
} E4hipStepConfig;

/* Library set-up on one GPU; creates the library's HIP stream.
 * Replaces nothing in the reference (the D runtime has no device set-up). */
int e4hip_init(int device);
int e4hip_finalize(void);
/* Copies the last error text (NUL-terminated, truncated to len). */
int e4hip_last_error(char* buf, int len);
/* The hipStream_t all work is queued on (for event timing by callers). */
void* e4hip_stream(void);
int e4hip_synchronize(void);

/* Gas models.  Ideal gas from molecular mass and gamma
 * (gas/ideal_gas.d:31-72, R = 8.31451/mMass). */
int e4hip_gas_ideal(double mMass, double gamma, int* gm);

/* Thermally perfect mixture (gas/therm_perf_gas.d:41-110): nsp species of
 * molecular masses M[nsp], each with nseg NASA-9 (CEA) coefficient sets
 * coeffs[nsp][nseg][9], break points Tbreaks[nsp][nseg+1] and blend ranges
 * Tblends[nsp][nseg-1] (gas/thermo/cea_thermo_curves.d:22-45).  The device path
 * is built for nsp = 5 (N2, O2, NO, N, O air). */
int e4hip_gas_tp(int nsp, const double* M, int nseg, const double* Tbreaks, const double* Tblends,
                 const double* coeffs, int* gm);

/* Learned initial guess of the thermally perfect temperature solve (no
 * counterpart in the reference; it replaces the starting point of
 * therm_perf_gas_mix_eos.d:70-153 in the decode).  A 7 -> 64 -> 64 -> 64 -> 1
 * SiLU network on (rho, u, Y0..Y4), evaluated in BF16 MFMA with FP32
 * accumulation, then two FP64 Newton steps accepted by the energy residual at
 * the result (when |residual / (dz/dT)| <= 1e-11 T); a cell that fails the
 * test falls back to the reference's bracketed solve from its stored
 * temperature.  params: E4HIP_MLP_NPARAM floats in the training layout
 * (W1[64][8], b1[64], W2[64][64], b2, W3[64][64], b3, w4[64], b4); features are
 * normalised as (x - in_mean) * in_std, and T = T_mean + T_std * output. */
#define E4HIP_MLP_NPARAM 8961
int e4hip_gas_tp_set_guess_model(int gm, const float* params, const float* in_mean, const float* in_std,
                                 float T_mean, float T_std);
/* Guesses for n states x[n][7] (rho, u, Y0..Y4) -> T[n] (host arrays). */
int e4hip_tp_guess(int gm, int n, const double* x, double* T);
/* Energy evaluations of the decode temperature solves, and the decodes whose
 * polish (mixed mode or learned guess) fell back to the bracketed solve, since
 * the last reset (summed over blocks and stages).  No counterpart in the
 * reference: work counts of the solves that replace therm_perf_gas_mix_eos.d:70-153. */
int e4hip_tp_decode_stats(long long* energy_evals, long long* fallbacks, int reset);
/* The same for the face-state solves of the reconstruction (two per face,
 * onedinterp.d:534-551); fallbacks count only in the mixed mode. */
int e4hip_tp_face_stats(long long* energy_evals, long long* fallbacks, int reset);
/* Training of the guess network with this library's own kernels (FP32).
 * X[n][8] normalised features (8th = 0), T[n] normalised targets.
 * e4hip_mlp_grad: mean squared error and its gradient for one batch.
 * e4hip_mlp_train: Adam with cosine learning-rate decay, params updated in place,
 * loss_per_epoch (may be NULL) receives [epochs] values. */
int e4hip_mlp_grad(const float* params, const float* X, const float* T, int n, float* loss, float* grads);
int e4hip_mlp_train(float* params, const float* X, const float* T, int n, int epochs, int batch, float lr,
                    unsigned long long seed, float* loss_per_epoch);

/* Blocks.  Uniform Cartesian cells of size dx*dy*dz (dz ignored for dims 2);
 * replaces SFluidBlock construction + compute_primary_cell_geometric_data
 * (eilmer/sfluidblock.d:48-126, 650-842) for axis-aligned grids.  Every face
 * starts as a slip wall (prep.lua:535, 551 default). Each active direction
 * needs at least 2 cells. */
int e4hip_block_create(int gm, int dims, int ni, int nj, int nk,
                       double dx, double dy, double dz, int* blk);
/* Blocks of any structured grid (compute_primary_cell_geometric_data,
 * eilmer/sfluidblock.d:650-842, as the D side has already computed it for its
 * SFluidBlock).  cartesian != 0: geom = {dx, dy, dz}, the same as
 * e4hip_block_create.  cartesian == 0: geom holds the cell geometry
 * [4][ncells] (volume -- the (x,y)-plane area in 2D --, iLength, jLength,
 * kLength; fvcell.d:58-154) and then, for each direction d < dims, the face
 * geometry [10][nfaces(d)] (area, n.x, n.y, n.z, t1.x, t1.y, t1.z, t2.x, t2.y,
 * t2.z; fvinterface.d:34-90), cells i fastest and faces of direction d i
 * fastest over (ni + [d==0]) x (nj + [d==1]) x (nk + [d==2]); d = 0 are the
 * i-faces (west faces of the cells, normal towards +i), d = 1 the j-faces and
 * d = 2 the k-faces.  The ghost cells' lengths are the reference's: mirrored
 * from the interior (sfluidblock.d:673-713), or across a full-face connection
 * the mapped cell's own (full_face_copy.d:1042-1085).  Such a block's stages
 * reconstruct with the four cells' lengths (onedinterp.d:118-162) in each
 * face's own frame (vector3.d:425-450, fluxcalc.d:26-128).  It takes every
 * boundary kind except E4HIP_BC_SLIP_WALL_1 (and, for an ideal gas,
 * E4HIP_BC_EXTRAPOLATE_1), and connects only to blocks also built this way.
 * Volumes and lengths must be positive. */
int e4hip_block_create_geometry(int gm, int dims, int ni, int nj, int nk, const double* geom, int cartesian,
                                int* blk);
int e4hip_block_destroy(int blk);
/* Boundary condition of one face (E4HIP_FACE_*, E4HIP_BC_*).  For
 * E4HIP_BC_INFLOW, fs gives the ghost flow state [8 + nsp] as above; for
 * E4HIP_BC_INFLOW_PROFILE, fs holds one state per ghost cell,
 * [2 layers][face cells][9 + nsp] = rho, p, T, u, a, vx, vy, vz, S, Y..., with
 * the face cells in-face index a + b * na (north/south: a = i, b = k;
 * east/west: a = j, b = k; top/bottom: a = i, b = j); else fs may be NULL.
 * A wall without ghost data (E4HIP_BC_SLIP_WALL_1) needs at least 3 cells
 * across it.  Periodic faces must be set on both opposite faces.
 * Replaces BoundaryCondition.applyPreReconAction for these effects
 * (eilmer/bc/boundary_condition.d:274-292) and the exchange of
 * simcore.d:1599-1668 for a block connected to itself. */
int e4hip_block_set_bc(int blk, int face, int type, const double* fs);
/* Full-face connection of face `face` of blk to face `other_face` of block
 * `other` with orientation 0..3 (0 in 2D): the ghost cells of the face copy
 * the mapped interior cells of the other block at the start of every stage.
 * One side of the pair, as one ExchangeBC_FullFace{otherBlock, otherFace,
 * orientation} (prep.lua:681-684, full_face_copy.d:142-962, exchange
 * simcore.d:1599-1668); connectBlocks sets both sides with the same
 * orientation.  The blocks must share the gas model and dimensionality and the
 * face sizes must match under the orientation.  blk == other is allowed. */
int e4hip_block_connect(int blk, int face, int other, int other_face, int orientation);
/* 2D axisymmetric geometry (GlobalConfig.axisymmetric, x the axis, y the
 * radius) for a block whose lower-left vertex is (x0, y0), y0 >= 0: per-radian
 * cell volumes (area * centroid y, fvcell.d:314-351; properties.d:177-237) and
 * face areas (length * mid-point y, fvinterface.d:265-297), and the pressure
 * source p * areaxy / volume of the y-momentum (add_inviscid_source_vector,
 * fvcell.d:2177-2181).  Planar blocks need no call. */
int e4hip_block_set_axisymmetric(int blk, double x0, double y0);
/* User-defined source terms of a block (GlobalConfig.udf_source_terms;
 * addUDFSourceTermsToCell, user_defined_source_terms.d:23-56, called from
 * simcore.d:2445-2464): Q is [nq][ncells] in the conserved order (mass, x/y/z
 * momentum, total energy, then the species densities of a mixture), added to
 * each cell's source vector after the axisymmetric term at every stage.  The
 * reference calls its Lua function at every stage with the cell's state; here
 * the values are held constant until the next call (a manufactured-solution
 * source, or a caller refreshing them between steps).  NULL removes them. */
int e4hip_block_set_source_terms(int blk, const double* Q);
/* Host -> device copy of U[0] ([nq][ncells]); the flow state is the decode of
 * U (fvcell.d:563-732).  For a thermally perfect gas, Tguess[ncells] is the
 * temperature the Newton solve starts from (the flow state's T on restart);
 * for an ideal gas it may be NULL.  Returns E4HIP_ERR_DECODE if some cell
 * does not decode to a valid state. */
int e4hip_block_upload(int blk, const double* U, const double* Tguess);
/* Device -> host copy of U[0] and (if fs != NULL) the decoded flow state. */
int e4hip_block_download(int blk, double* U, double* fs);
int e4hip_block_ncells(int blk, long* ncells, int* nq);

/* Configuration shared by all blocks (GlobalConfig/LocalConfig fields). */
int e4hip_config(const E4hipStepConfig* c);

/* gasdynamic_explicit_increment_with_fixed_grid() (simcore.d:2303-2952) for
 * the listed blocks and time step dt: every stage of the configured scheme
 * (stage s writes U[s] from U[s-1]), ending with U[0] holding the new state
 * (the swap of U[0] and U[stages], simcore.d:2936-2941).
 * n_bad_cells (may be NULL) receives, per stage (3 ints), the invalid-cell
 * count (count_invalid_cells, fluidblock.d:433-499) plus failed decodes,
 * summed over the blocks.  As in the reference, the counts are checked after
 * each stage: when a block's count exceeds max_invalid_cells the later stages
 * do nothing and E4HIP_ERR_BAD_CELLS is returned (simcore.d:2480-2493,
 * 2694-2696, 2897-2899); a decode that would throw (non-positive density)
 * gives E4HIP_ERR_DECODE the same way.  Either way U[0] (and a mixture's flow
 * state) is left as it was before the call, and the caller ends the run. */
int e4hip_explicit_increment(const int* blks, int nblk, double dt, int* n_bad_cells);
/* The same counts per block: counts[n][s] for the n-th block of the last
 * e4hip_explicit_increment call (nblk must be that call's) and stage s < 3,
 * 0 for stages that did not run -- the per-block figures the reference prints
 * before it throws (simcore.d:2480-2487).  The caller advances the simulation
 * time (SimState.time = t0 + dt, simcore.d:2951) after a successful call: no
 * kernel of the path depends on t. */
int e4hip_last_step_bad_cells(int nblk, int* counts);

/* determine_time_step_size for the listed blocks (fvcell.d:1624-1705;
 * fluidblock.d:776-862): the CFL signal of each cell is the non-stringent one
 * or, with E4hipStepConfig.stringent_cfl, the fastest normal velocity over the
 * shortest cell length (fvcell.d:1642-1650); dt_allow = min over cells of
 * cfl_value/signal and cfl_max = max over cells of dt_current*signal,
 * reduced over blocks (simcore.d:1385-1394), with the reference's rescaling
 * when check_cfl and cfl_max exceeds cfl_allow. */
int e4hip_determine_dt(const int* blks, int nblk, double dt_current, double cfl_value,
                       int check_cfl, double* dt_allow, double* cfl_max);

/* e4hip_determine_dt without device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same inputs, outputs
 * and results as e4hip_determine_dt (bit for bit); its per-block signals and
 * the reduction live in a fixed device array of the library (one per GPU, held
 * until the process ends), and the whole call synchronises once.  At most
 * E4HIP_NOALLOC_MAX_BLOCKS blocks (E4HIP_ERR_ARG beyond).  Calls that overlap
 * on one GPU (several host threads, or with each other's work still queued on
 * another stream) race on that array. */
#define E4HIP_NOALLOC_MAX_BLOCKS 4096
int e4hip_determine_dt_noalloc(const int* blks, int nblk, double dt_current, double cfl_value,
                               int check_cfl, double* dt_allow, double* cfl_max);

/* Per-kernel timing with HIP events on the library stream (off by default).
 * e4hip_profile_read returns, for kernel class k (0 = stage kernels,
 * 1 = CFL kernel, 2 = the decode of an uploaded mixture state), the summed
 * milliseconds and launch count since the last read, and resets them. */
int e4hip_profile_enable(int on);
int e4hip_profile_read(int k, double* total_ms, long* launches);

/* The k-chunk height (planes per workgroup) the stage kernel of a uniform 3D
 * ideal-gas block launches with on its own under the current configuration:
 * the autotuner's choice once its shape was tuned, else the default.
 * No counterpart in the reference (its loops have no tiling). */
int e4hip_block_kchunk(int blk, int* kchunk);

#ifdef __cplusplus
}
#endif
#endif /* E4HIP_H */
