/*
 * bubbledynamics.h -- C ABI of the MI355X (gfx950) focused-ultrasound kernels.
 *
 * Every entry point takes plain device pointers, plain integers and a HIP
 * stream passed as `void*` (NULL = the default stream), and returns an int
 * status (BD_OK = 0).  No C++ or torch types cross this boundary.  Buffers are
 * owned by the caller; a plan object owns only its scratch (reduction
 * partials, step counter, time table, checkpoints).
 *
 * Drop-in entry points mirror one reference kernel launch each, with the same
 * grid integers and argument meaning; the reference's __constant__/__device__
 * symbols become an explicit parameter struct.  Reference files:
 *   W = Westervelt_relax_fft_Pressure_Energy.cu
 *   T = Westervelt_Temperature_SquareSpiral.cu
 *   G = Gilmore_liver_3fr_combo.cu
 *
 * Field layout (all 3-D fields): idx = iz*Nxy*Nxy + iy*Nxy + ix, x fastest
 * (W:69, T:59).  dtype selects the element type of every field argument:
 * BD_F32 (float) or BD_F64 (double).
 */
#ifndef BUBBLEDYNAMICS_H
#define BUBBLEDYNAMICS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BD_OK               0
#define BD_ERR_INVALID      1  /* bad argument (null pointer, size, dtype)      */
#define BD_ERR_UNSTABLE     2  /* parameters violate the scheme's stability    */
#define BD_ERR_ALLOC        3  /* device allocation failed or exceeds the cap  */
#define BD_ERR_HIP          4  /* a HIP runtime call failed                    */
#define BD_ERR_NONFINITE    5  /* a reduction produced NaN/Inf                  */
#define BD_ERR_UNSUPPORTED  6  /* option combination not implemented           */
#define BD_ERR_IO           7  /* a file could not be opened, read or written   */

#define BD_F32 0
#define BD_F64 1

/* ======================================================================== */
/*  Westervelt acoustics                                                    */
/* ======================================================================== */

#define BD_BC_MUR   0   /* first-order Mur on 3 layers (W:88-135), outward-ordered */
#define BD_BC_ZERO  1   /* boundary layers held at p = 0                            */

#define BD_SRC_CAP    0 /* phase-delayed spherical cap with hole (W:62-85)          */
#define BD_SRC_ARRAY  1 /* square tiling of piston elements, one (A, phi) each      */

/*
 * Parameters of one Westervelt run.  The first block is the reference's
 * __constant__ block (W:16-52) field for field.  The second block holds the
 * extensions; all zeros reproduce the reference.
 */
typedef struct bd_westervelt_params {
    /* water region (z < water_tissue), W:16-20 */
    double rho, c_inf, delta, beta, abs_water;
    /* tissue region (z >= water_tissue), W:23-33 */
    double rho_tissue, c_inf_tissue, beta_tissue, abs_tissue;
    double delta_tissue, c1, c2, tau1, tau2;
    /* transducer, W:37-42 */
    double fr, d, alpha, alpha_hole, p0;
    double water_tissue;               /* W:44 */
    /* mesh, W:48-52 */
    double xy_min, dxy, z_min, dz, dt;

    /* ---- extensions (0 = reference behaviour) ---- */
    int32_t full_domain;   /* 0: quarter domain, even mirror at ix=0, iy=0 (W:176-200); 1: full domain */
    int32_t boundary;      /* BD_BC_MUR or BD_BC_ZERO                                               */
    int32_t source;        /* BD_SRC_CAP or BD_SRC_ARRAY                                            */
    int32_t sponge_cells;  /* damping layer width at the outer faces (0 = none)                     */
    double  sponge_alpha;  /* per-step factor exp(-(sponge_alpha*(W-dist))^2) inside the layer      */
    double  burst_cycles;  /* 0: continuous wave (W:82); >0: rectangular gate of that many cycles   */
    double  burst_t0;      /* gate opens at delayed time tt - burst_t0 = 0                          */
    /* BD_SRC_ARRAY: elements tile [arr_x0, arr_x0+arr_nx*arr_pitch) x [arr_y0, ...) on plane iz=2 */
    int32_t arr_nx, arr_ny;
    double  arr_x0, arr_y0, arr_pitch;
    double  reserved[8];
} bd_westervelt_params;

/* Drop-in for Westervelt_space6<<<>>> (W:139-236): interior update of Pnext
 * and of the relaxation variables R1, R2 (tissue region only). */
int bd_westervelt_space6(const bd_westervelt_params* prm, int dtype,
                         void* Pnext, const void* P, const void* Pm1, const void* Pm2,
                         const void* Pm3, const void* Pm4,
                         void* R1, const void* R1m1, const void* R1m2,
                         void* R2, const void* R2m1, const void* R2m2,
                         int Nxy, int Nz, void* stream);

/* Drop-in for boundary_condition_source<<<>>> (W:62-85): hard source on
 * plane iz=2 at time t (t = t^{n+1} as in W:564-566).  For BD_SRC_ARRAY,
 * elem_amp/elem_phase hold arr_nx*arr_ny doubles each (device memory);
 * they are ignored (may be NULL) for BD_SRC_CAP. */
int bd_boundary_condition_source(const bd_westervelt_params* prm, int dtype, void* P,
                                 double t, int Nxy, int Nz,
                                 const double* elem_amp, const double* elem_phase,
                                 void* stream);

/* Drop-in for boundary_condition_space6<<<>>> (W:88-135): first-order Mur
 * faces (z-min, z-max, x-max, y-max; plus x-min, y-min in a full domain),
 * evaluated in a defined order: z faces, then x faces, then y faces, each
 * face innermost layer first, every layer reading its neighbour's new value.
 * P = Pnext (source plane already set), Pm1 = p^n. */
int bd_boundary_condition_space6(const bd_westervelt_params* prm, int dtype,
                                 void* P, const void* Pm1, double t, int Nxy, int Nz,
                                 void* stream);

/* Drop-in for acoustic_power_water<<<>>> (W:242-265).  last=0 accumulates
 * Q += ((Pnext-Pm1)/(2dt))^2 dt in the water region; last=1 applies the
 * final scale 2*abs_water*fr/((2 pi fr)^2 rho c). */
int bd_acoustic_power_water(const bd_westervelt_params* prm, int dtype, void* Q,
                            const void* Pnext, const void* Pm1, int Nxy, int Nz,
                            int last, void* stream);

/* Replaces copy_data + Double_to_Complex + cufftExecC2C + Abs_Complex
 * (W:268-321, W:740-747) by running DFT accumulators: for harmonic n=1..N_harm,
 * H[2(n-1)]   += p * cos(2 pi n*periods*i_time/N_time),
 * H[2(n-1)+1] -= p * sin(2 pi n*periods*i_time/N_time),
 * on planes iz in [iz_tissue, Nz) (all ix, iy).  H holds 2*N_harm fields of
 * Nxy*Nxy*(Nz-iz_tissue) elements (same dtype); with kahan=1 it holds 4*N_harm
 * fields, the last 2*N_harm being Kahan compensations.  These are exactly the
 * FFT bins k = n*periods the reference reads (W:341). */
int bd_harmonics_accumulate(int dtype, void* H, const void* Pnext, int Nxy, int Nz,
                            int iz_tissue, int i_time, int N_time, int periods,
                            int N_harm, int kahan, void* stream);

/* Drop-in for total_energy_Nharmonics_xyz (W:324-345) over the whole tissue
 * slab: Q = sum_n 2*(abs_coef*n*fr/1e6) * 2*|X_n/N_time|^2/(rho_t c_t).
 * abs_coef = 7.2 reproduces W:342.  kahan as in bd_harmonics_accumulate. */
int bd_harmonics_q(const bd_westervelt_params* prm, int dtype, void* Q, const void* H,
                   int Nxy, int Nz, int iz_tissue, int N_time, int N_harm,
                   double abs_coef, int kahan, void* stream);

/* ---- run-level plan: pointer ring + fused kernels + hipGraph ----------- */
typedef struct bd_westervelt_plan bd_westervelt_plan;

/* State block: 12 fields (6 pressure levels, 3 R1 levels, 3 R2 levels)
 * laid out back to back, 12*Nxy*Nxy*Nz elements of the plan's dtype. */
size_t bd_westervelt_state_elems(int Nxy, int Nz);

/* max_steps bounds the per-step outputs (energy, focus series) and the
 * time table.  focus_* name the cell recorded every step (-1 = none). */
int bd_westervelt_plan_create(const bd_westervelt_params* prm, int dtype, int Nxy, int Nz,
                              int max_steps, int focus_ix, int focus_iy, int focus_iz,
                              bd_westervelt_plan** out);
/* Binds caller-owned device buffers: state (bd_westervelt_state_elems),
 * pmax (Nxy*Nxy*Nz, may be NULL), energy (2*max_steps doubles: per step
 * sum p^2 dV/(rho c^2) [J] and the leapfrog invariant, may be NULL),
 * focus (max_steps doubles, may be NULL), array element amp/phase (device,
 * arr_nx*arr_ny doubles each, BD_SRC_ARRAY only). */
int bd_westervelt_plan_bind(bd_westervelt_plan* plan, void* state, void* pmax, double* energy,
                            double* focus, const double* elem_amp, const double* elem_phase);
/* Advances n_steps steps (interior + source + ABC each step).  use_graph=1
 * replays a captured 6-step hipGraph (the ring period).  pmax is the running
 * maximum over every level this plan produced once the call's work on the
 * stream is done (inside a call the interior cells are updated once every 6
 * steps over the six levels a step holds, and the call's last step also takes
 * the levels after the last such window); levels present before a call (e.g.
 * restored after bd_westervelt_plan_seek) are not taken into pmax. */
int bd_westervelt_plan_run(bd_westervelt_plan* plan, int n_steps, int use_graph, void* stream);
/* Chooses the step kernel's tile by measurement, then captures and
 * instantiates the plan's pair of 6-step hipGraphs (after bind; a later bind
 * drops the graphs), so that no later bd_westervelt_plan_run with
 * use_graph=1 spends host time on graph construction.  The tile choice runs
 * once, before the plan's first step: each candidate tile relaunches the
 * next step a few times on a private stream (writing only what that step
 * will overwrite; pmax untouched) and the fastest is kept -- results do not
 * depend on the tile beyond FP rounding order of the energy partials.
 * Optional: without it the plan keeps a static tile and run builds the
 * graphs on first use.  Synchronous. */
int bd_westervelt_plan_prepare(bd_westervelt_plan* plan);
/* Reduces the per-block energy partials of all steps run so far into the
 * bound energy buffer. */
int bd_westervelt_plan_finalize(bd_westervelt_plan* plan, void* stream);
/* Steps run so far, and the ring slot holding p^{n-level} (level 0 = newest)
 * or R1/R2 levels (which = 0: P, 1: R1, 2: R2). */
int bd_westervelt_plan_steps(const bd_westervelt_plan* plan);
/* First plane iz with z_min + iz*dz >= water_tissue (the medium switch of
 * W:212; the reference's iz_tissue = 1202, W:674), Nz if none, -1 on bad
 * arguments. */
int bd_westervelt_first_tissue_plane(const bd_westervelt_params* prm, int Nz);

/* Declares that the bound state holds the levels of step `step` (0..max_steps)
 * in their ring slots, as a plan run to that step leaves them -- e.g. a copy
 * of such a state restored into the bound block; the next run continues at
 * t = t^step (the restart of W:663-704). */
int bd_westervelt_plan_seek(bd_westervelt_plan* plan, int step);
int bd_westervelt_plan_slot(const bd_westervelt_plan* plan, int which, int level);
int bd_westervelt_plan_destroy(bd_westervelt_plan* plan);
/* Launches the step kernel of the next step again (same arguments as the
 * real step, state not advanced; it rewrites the next levels and may raise
 * pmax) -- used to time that kernel in isolation.  Requires steps < max_steps. */
int bd_westervelt_plan_rerun_step(bd_westervelt_plan* plan, void* stream);
/* Number of inside cells (the 6th-order update) per step; the other cells of
 * the grid are Mur faces and the source plane. */
long long bd_westervelt_plan_interior_cells(const bd_westervelt_plan* plan);
/* Step-kernel launch shape chosen for the plan: tile columns, thread rows,
 * rows per thread, planes of loads in flight (0-3), history-load cache policy (0 cached,
 * 1 every level non-temporal, 2 only p^{n-3}, p^{n-4}, R^{n-2} non-temporal) and
 * workgroups per launch.  Any output pointer may be NULL.  Introspection
 * only; the reference has one fixed launch shape (W:560). */
int bd_westervelt_plan_shape(const bd_westervelt_plan* plan, int* tile_x, int* tile_y, int* rows,
                             int* prefetch, int* nontemporal, int* workgroups);

/* ======================================================================== */
/*  Bioheat (heat conduction + acoustic source, CEM43 thermal dose)         */
/* ======================================================================== */

/* The thermal part of the reference's __constant__ block (T:35-45) plus the
 * Pennes extension; perf = 0 and T_off = 0 reproduce the reference. */
typedef struct bd_bioheat_params {
    double water_tissue;        /* T:30: planes with iz*dz < water_tissue use the water set */
    double dxy, dz, dt_T;       /* T:35-37 */
    double kdt_dxy, kdt_dz, kdt_dxy_water, kdt_dz_water;   /* T:40-43: k dt / h^2 */
    double dt_rho_ct, dt_rho_ct_water;                     /* T:44-45: dt / (rho c_t) */
    double perf;                /* Pennes: w_b c_b dt / (rho c_t) in tissue (0 = none) */
    double T_art;               /* arterial temperature [K] */
    double T_off;               /* extended step only: fields hold T - T_off [K] */
    double reserved[8];
} bd_bioheat_params;

/* Drop-in for Temperature_square<<<>>> (T:52-105): 7-point explicit update of
 * the interior, shifted acoustic source Q[ix-shift_x, iy-shift_y, iz] (0
 * outside), CEM43 dose TD += R^(43-T_c) dt_T.  use_p=1 reads T_p and writes
 * T_n; use_p=0 the reverse. */
int bd_temperature_square(const bd_bioheat_params* prm, int dtype, void* T_n, void* T_p,
                          const void* Q, void* TD, int Nxy, int Nz, int use_p,
                          int shift_x, int shift_y, void* stream);

/* Extended step: T_out <- update(T_in), Pennes perfusion, fields stored as
 * T - T_off, and with TD_comp != NULL a Kahan-compensated dose sum. */
int bd_bioheat_step(const bd_bioheat_params* prm, int dtype, void* T_out, const void* T_in,
                    const void* Q, void* TD, void* TD_comp, int Nxy, int Nz,
                    int shift_x, int shift_y, void* stream);

/* n_steps substeps of bd_bioheat_step at one source shift, alternating
 * between the fields T_a and T_b: *which (0 or 1) names the field holding the
 * current temperature on entry and on return.  Asynchronous on the stream.
 * The boundary cells are held fixed: on entry the other field's boundary
 * shell is set to the current field's (the reference initialises both fields
 * alike, T:233-257).  Pairs of substeps run in one launch (bit-identical
 * results; each pair flips *which once, so the other field then holds the
 * level two substeps back). */
int bd_bioheat_run(const bd_bioheat_params* prm, int dtype, void* T_a, void* T_b, const void* Q,
                   void* TD, void* TD_comp, int Nxy, int Nz, int shift_x, int shift_y, int n_steps,
                   int* which, void* stream);

/* The reference's continuous square-spiral motion (T:141-173, T:323-409):
 * per time step the integer source shift (shift_x, shift_y).  start is the
 * initial focus coordinate (x = y = start; the reference uses
 * -0.085 + 425*dxy, T:331-332).  Writes at most cap entries and returns the
 * number of steps (77,474 for the reference's 13 lines), or -1 on bad
 * arguments. */
int bd_spiral_reference(double speed, double scale, double speed_init, double t_acc,
                        double dt, double dxy, double start, int lines, int* shift_x,
                        int* shift_y, int cap);

/* Dwell spiral: n_side x n_side focus positions at pitch (m) on the grid
 * spacing dxy, visited as a square spiral from the centre (legs 1,1,2,2,...
 * with the reference's leg order: +x, +y, -x, -y, ...).  Writes n_side^2
 * integer shifts (rounded to the nearest cell). */
int bd_spiral_dwell(int n_side, double pitch, double dxy, int* shift_x, int* shift_y);

/* Writes a slice of a device field in the reference's text format
 * (T:412-447): header, then one line per row with every value followed by
 * ';' (printf %g).  BD_SLICE_XZ: rows iz in [r0,r1), columns ix in [c0,c1) at
 * iy = fixed, header "Z=rows;X=columns" (the reference: iz 1200..1700,
 * ix 325..525, iy 425); BD_SLICE_XY: rows iy, columns ix at iz = fixed,
 * header "Y=rows;X=columns" (iz 1500).  offset is added to every value (a
 * field stored as T - T_off).  Synchronous. */
#define BD_SLICE_XZ 0
#define BD_SLICE_XY 1
int bd_write_slice(const char* path, int dtype, const void* field, int Nxy, int Nz, int plane, int fixed,
                   int r0, int r1, int c0, int c1, double offset, void* stream);

/* ---- Westervelt -> bioheat hand-off ------------------------------------- */

/* Writes the heat source in the reference's Q3D file format (W:778-790):
 * planes iz0..Nz-1 of the (quarter-domain) field Q (device, Nxy x Nxy x Nz),
 * as float64, x fastest, then y, then z; the reference writes iz0 = 2 (the
 * plane z = 0).  FP32 fields are widened.  Synchronous. */
int bd_q3d_write(const char* path, int dtype, const void* Q, int Nxy, int Nz, int iz0, void* stream);

/* Reads a Q3D file as the reference's thermal program does (T:271-302):
 * (Nxy_full/2+1)^2 * Nz float64 values of a quarter domain, expanded to the
 * full Nxy_full x Nxy_full x Nz grid (device, dtype) by the mirror
 * (ix, iy) -> (|ix - Nxy_full/2|, |iy - Nxy_full/2|).  BD_ERR_IO if the file is
 * missing or shorter.  Allocates a staging buffer; synchronous. */
int bd_q3d_read_expand(const char* path, int dtype, void* Q_full, int Nxy_full, int Nz, void* stream);

/* The same mirror expansion between device buffers: Q_quarter holds
 * (Nxy_full/2+1)^2 * Nz values of the dtype. */
int bd_mirror_expand(int dtype, const void* Q_quarter, void* Q_full, int Nxy_full, int Nz, void* stream);

/* ======================================================================== */
/*  Gilmore bubble dynamics in viscoelastic tissue                          */
/* ======================================================================== */

/* The reference's __device__ symbols (G:13-33) plus the integrator's
 * controls.  bd_gilmore_reference_params() fills the liver values of G:440-460. */
typedef struct bd_gilmore_params {
    double rho, P0, S2, C_inf, lambda, n, B, gamma1, gamma3, mu4, G4_3;  /* G:13-23 */
    int32_t triple_fr;                                                   /* G:26    */
    int32_t tableau;           /* BD_GM_TAB_REFERENCE (G:218-227 as written, a32 = -9/40)
                                  or BD_GM_TAB_DP1980 (published Dormand-Prince, a32 = +9/40) */
    double minA, stepA, minR0, stepR0;                                   /* G:29-33 */
    double abs_eps, rel_eps;   /* DP5(4) tolerances, G:214-215 (1e-9, 1e-9)       */
    double h0;                 /* first step of every solve, G:355 / G:393 (1e-13) */
    double max_solve_steps;    /* cap on accepted steps of one ODE solve (0 = 1e7); the
                                  reference has none and can spin forever once h underflows */
    int32_t arithmetic;        /* BD_GM_ARITH_REFERENCE or BD_GM_ARITH_FAST                    */
    int32_t pad0;
    double reserved[6];
} bd_gilmore_params;

#define BD_GM_TAB_REFERENCE 0
#define BD_GM_TAB_DP1980    1

/* Arithmetic of the right-hand side and of the first Runge-Kutta stage.
 * REFERENCE evaluates G:40-179 in the reference's expression order (the
 * cancelling forms pow(I, (n-1)/n) - 1 and (p_in - 2s/R + tau + B)/(P0 + p + B)
 * included) and recomputes the first stage of every step (G:259-273), so
 * results follow the reference to the last bits its libm leaves open; it is
 * the default of bd_gilmore_reference_params.  FAST shares subexpressions,
 * forms I^((n-1)/n) - 1 without cancellation and reuses the last stage of a
 * step as the next step's first (FSAL): 6 instead of 7 right-hand sides per
 * attempt, about 3x faster, equal in exact arithmetic and not bit-identical. */
#define BD_GM_ARITH_REFERENCE 0
#define BD_GM_ARITH_FAST      1

/* Threshold-sweep result of a cell whose search met an ODE solve that ended
 * neither in collapse nor at endT within max_solve_steps accepted steps (the
 * reference, which has no cap, would have kept integrating that solve). */
#define BD_GM_SWEEP_CAPPED (-2.0)

void bd_gilmore_reference_params(bd_gilmore_params* p);

/* Drop-in for DPRK_solver_threshold_fix_tissue<<<>>> (G:323-429): one result
 * per (frequency combo, R0) pair, idx = i_combo*N_R0 + i_R0; the inertial-
 * cavitation threshold amplitude found in 0.1 / 0.01 / 0.001 MPa stages, -1
 * if none is set, 1e7 if no collapse up to 10 MPa, BD_GM_SWEEP_CAPPED if a
 * solve reached max_solve_steps.  Fr_combo: 3*N_fr_combo frequencies
 * (device); at most 2^31 - 2^24 - 1 cells (32-bit queue counters).  N_A is
 * accepted and ignored, as in the reference.  The queue
 * counter is one per device: sweeps that overlap in time on one device (other
 * streams or host threads) race; run them one after another. */
int bd_gilmore_threshold_sweep(const bd_gilmore_params* prm, double* results_A,
                               const double* Fr_combo, int N_fr_combo, int N_R0, int N_A,
                               void* stream);

#define BD_GM_RK4   0   /* classical fixed-step Runge-Kutta 4                       */
#define BD_GM_DP45  1   /* adaptive Dormand-Prince 5(4), controller of G:212-318    */

#define BD_GM_FP64   0  /* all arithmetic in FP64                                    */
#define BD_GM_SPLIT  1  /* FP64 state/error norm; FP32 right-hand side while every bubble of the
                           wavefront has R/R0 >= split_ratio (results then depend on which
                           bubbles share a wavefront, i.e. on the launch shape)          */

/* One batch of independent bubbles driven by
 * p(t) = sum_k amp[3i+k] sin(2 pi freq[k] t + phase[3i+k]).
 * A bubble stops at t_end or at collapse (dR/dt <= -C_inf, the criterion of
 * G:313).  All arrays are device memory. */
typedef struct bd_gilmore_batch {
    int32_t method, precision, n;
    int32_t out_every;        /* RK4: record R every out_every steps (0 = no record)   */
    int32_t n_out;            /* samples per bubble in R_out (NaN after collapse)      */
    int32_t block;            /* workgroup size (0 = tuned default)                    */
    int32_t chunk;            /* bubbles taken per queue grab (0 = tuned default)      */
    int32_t waves;            /* wavefronts per SIMD of the persistent grid (0 = tuned default) */
    double freq[3];
    double t_end, dt;         /* dt: RK4 step, DP45 first step                         */
    double rtol, atol;        /* DP45 tolerances                                        */
    double split_ratio;       /* BD_GM_SPLIT switches to FP64 RHS below R/R0 = ratio   */
    long long max_steps;      /* per-bubble cap on attempted steps (0 = 1e7)           */
    const double* R0; const double* amp; const double* phase;
    double* R_out;            /* n * n_out, nullable                                   */
    double* Rmax;             /* n                                                       */
    double* t_collapse;       /* n, -1 if no collapse before t_end                       */
    long long* n_acc;         /* n, accepted steps                                      */
    long long* n_rej;         /* n, rejected attempts (DP45)                            */
    int32_t schedule;         /* BD_GM_SCHED_*: order in which the queue hands out bubbles */
    int32_t slice;            /* BD_GM_SCHED_SLICED: attempts of the first pass (0 = 128) */
    double reserved[3];
} bd_gilmore_batch;

/* Allocation-free variant of bd_gilmore_threshold_sweep (no counterpart in
 * the reference): same arguments, results and kernels, but the sweep's queue
 * counters live in a static device array of the library (256 bytes, one per
 * GPU, kept until the process ends) instead of memory the library allocates
 * on a device's first call, so it never calls hipMalloc -- for callers that
 * cannot allocate (stream capture into a graph, serving loops that forbid it).
 * At most 2^31 - 2^24 - 1 cells (N_fr_combo * N_R0); more return
 * BD_ERR_INVALID with a message on stderr.  Every call on a GPU shares that
 * array: calls that overlap in time on one GPU race; run them one after
 * another. */
int bd_gilmore_threshold_sweep_noalloc(const bd_gilmore_params* prm, double* results_A,
                                       const double* Fr_combo, int N_fr_combo, int N_R0, int N_A,
                                       void* stream);

/* The same sweep over the cells idx in [cell0, cell_end) only, for runs too
 * long for one sitting (no counterpart in the reference; the full 27,775 x 100
 * table).  Cells are claimed in increasing idx and every claimed cell is
 * completed, so the cells done always form a prefix [cell0, cell0 + done).
 * While it runs the kernel writes its progress to the host-mapped words of
 * bd_gilmore_sweep_flags(): [0] cells claimed, [1] cells completed (both
 * approximate while running), and after the launch [1] = the exact count and
 * [3] = 1.  Setting [2] to nonzero asks the kernel to stop claiming cells;
 * the call leaves [2] as the caller set it, so clear it before a new call.
 * One range sweep at a time per process; call again only after [3] is 1. */
int bd_gilmore_threshold_sweep_range(const bd_gilmore_params* prm, double* results_A,
                                     const double* Fr_combo, int N_fr_combo, int N_R0, int cell0,
                                     int cell_end, void* stream);
/* The host-mapped progress / stop words of bd_gilmore_threshold_sweep_range
 * (4 ints, allocated on first use, kept until the process ends); NULL on
 * allocation failure. */
int* bd_gilmore_sweep_flags(void);

/* ---- resumable threshold sweep (no counterpart in the reference) --------
 * The drop-in sweep over cells [cell0, cell_end) in launches of bounded
 * length, for tables whose cells take hours (the reference tableau at low
 * frequency).  Each bd_gilmore_sweep_advance is one persistent launch in
 * which every lane stops after `budget` DP5(4) attempts (0 = no limit) or
 * within 256 attempts of a stop request: a lane holding an unfinished cell
 * saves its whole search state (stage, amplitudes, thresholds, ODE state,
 * first stage, step size, controller, step count) and the next advance
 * continues those cells before it claims new ones.  Results equal the
 * unsliced drop-in's bit for bit.  The handle owns its device state (about
 * 150 bytes per lane of the persistent grid) and host-mapped progress words. */
typedef struct bd_gilmore_sweep bd_gilmore_sweep;
/* Fr_combo (device, 3*N_fr_combo doubles) must stay valid until destroy. */
int bd_gilmore_sweep_create(const bd_gilmore_params* prm, const double* Fr_combo, int N_fr_combo, int N_R0,
                            int cell0, int cell_end, bd_gilmore_sweep** out);
/* One launch on the stream (asynchronous).  results_A (device) has
 * N_fr_combo*N_R0 doubles indexed like the drop-in's; a cell's entry is
 * written once, when its search ends. */
int bd_gilmore_sweep_advance(bd_gilmore_sweep* sw, double* results_A, long long budget, void* stream);
/* Progress after the last advance (read from host-mapped words; exact once
 * out[4] is 1): out[0] cells claimed, out[1] cells completed, out[2] cells
 * suspended (continued by the next advance), out[3] 1 when every cell of the
 * range is completed, out[4] 1 when the last advance's launch has ended. */
int bd_gilmore_sweep_progress(const bd_gilmore_sweep* sw, long long out[5]);
/* Host-mapped stop word of the handle: nonzero makes the running and later
 * advances claim nothing new and suspend their cells; clear it to go on. */
int* bd_gilmore_sweep_stop_word(bd_gilmore_sweep* sw);
int bd_gilmore_sweep_destroy(bd_gilmore_sweep* sw);

/* Queue order of bd_gilmore_integrate (the results do not depend on it):
 * INPUT hands bubbles out in input order.  SLICED runs passes of `slice`,
 * 16*slice and then unlimited attempts per bubble: a bubble still running at
 * the end of a slice is suspended with its state saved, and the next pass
 * continues the suspended bubbles ordered by their extrapolated remaining
 * attempts (quarter-octave buckets, most expensive first), so the longest
 * bubbles do not start late and hold up the end of the run.  SLICED keeps a
 * per-device scratch (about 160 bytes per bubble) until the process ends and
 * synchronises the stream between passes; SLICED calls that overlap in time
 * on one device share that scratch and race, so run them one after another. */
#define BD_GM_SCHED_INPUT  0
#define BD_GM_SCHED_SLICED 1

/* Bytes of device workspace bd_gilmore_integrate needs (queue counter). */
size_t bd_gilmore_workspace_bytes(void);
/* Integrates a batch with a persistent work-queue kernel: each lane takes the
 * next bubble from a device counter as soon as its previous one finished. */
int bd_gilmore_integrate(const bd_gilmore_params* prm, const bd_gilmore_batch* batch,
                         void* workspace, void* stream);

/* Writes a threshold table (host arrays: results[combo * n_r0 + i_r0] as
 * bd_gilmore_threshold_sweep returns it) in the reference's output format
 * (G:580-606): header "fr1;fr2;fr3 --> R0;R0...", one ';'-separated line
 * per triplet, closing line "time (s);seconds"; numbers as printf %g. */
int bd_gilmore_sweep_write(const char* path, const double* fr_combo, int n_combo, const double* results,
                           int n_r0, double min_r0, double step_r0, double seconds);
/* Autotuner for the launch shape of bd_gilmore_integrate (no counterpart in
 * the reference): times block in {64,128,256} x waves per SIMD in {2,4,8}
 * (and bubbles per grab in {1,4} at 8 waves) on the first `sample` bubbles of
 * the batch (0 = up to 2^20), each capped at probe_steps attempts (0 = 500),
 * and records the fastest for this device, method, precision and arithmetic; later
 * integrate calls use it for block / chunk / waves left at 0.  Returns the
 * choice and its rate (bubble-steps/s).  Allocates scratch; synchronous. */
int bd_gilmore_autotune(const bd_gilmore_params* prm, const bd_gilmore_batch* batch, int sample,
                        long long probe_steps, void* stream, int32_t* block, int32_t* chunk,
                        int32_t* waves, double* rate);

/* ---- adjoint: phased-array gradient (no reference counterpart) ---------- */

/* Objective J = -w_target sum_{n=1..Nt} sum_{target} p^2 + w_outside sum_n sum_{outside} p^2
 * over interior cells; the target is the sphere |cell - (cx,cy,cz)| <= radius (cells). */
typedef struct bd_adjoint_objective {
    int32_t cx, cy, cz, pad0;
    double radius, w_target, w_outside;
    double reserved[4];
} bd_adjoint_objective;

/* Runs n_steps forward steps of the full-domain phased-array model
 * (full_domain=1, BD_BC_ZERO, BD_SRC_ARRAY, one medium) from rest, returns J
 * (host) and, when grad_amp/grad_phase are non-NULL (device, arr_nx*arr_ny
 * doubles each), dJ/dA_e and dJ/dphi_e from the discrete adjoint.  Forward
 * levels are checkpointed every checkpoint_every steps and recomputed per
 * segment (checkpoint_every >= n_steps keeps them all); the recompute of a
 * segment runs on a second stream while the adjoint sweeps the segment after
 * it.  timings_ms (host, nullable) receives three wall times: the forward
 * pass, the part of the backward phase outside the adjoint sweep (recompute
 * not hidden under it), and the adjoint sweep.  Synchronous. */
int bd_westervelt_adjoint(const bd_westervelt_params* prm, int dtype, int Nxy, int Nz, int n_steps,
                          int checkpoint_every, const double* elem_amp, const double* elem_phase,
                          const bd_adjoint_objective* obj, double* J, double* grad_amp,
                          double* grad_phase, double* timings_ms, void* stream);

/* Library metadata */
const char* bd_version(void);
int bd_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* BUBBLEDYNAMICS_H */
