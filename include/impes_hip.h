/*
 * impes_hip.h — C ABI of the MI355X (gfx950) IMPES pressure/transport library.
 *
 * Conventions
 *   - every function returns 0 on success and a negative code on failure;
 *     impes_last_error() returns the message of the calling thread's last failure;
 *   - pointers named like arrays are DEVICE pointers owned by the caller (the
 *     library only borrows them for the call), except where marked "host";
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream); all work
 *     is enqueued on it; functions that return host values synchronise it;
 *   - indices are int32, values double (FP64) or float (FP32);
 *   - opaque handles own their device memory until *_destroy.
 *
 * Each entry point cites the reference interface (researchim-ai/oil-simulator-py)
 * whose role it takes over.
 */
#ifndef IMPES_HIP_H
#define IMPES_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define IMPES_ABI_VERSION 2

typedef struct impes_amg impes_amg;
typedef struct impes_solver impes_solver;
typedef struct impes_model impes_model;

const char* impes_last_error(void);
int impes_abi_version(void);

/* ------------------------------------------------------------------ operators */

/* y = A x (CSR). Replaces ClassicalAMG._spmv_csr(A, x, transpose=False),
 * src/solver/classical_amg.py:976-1009 (gather + scatter_add per call). */
int impes_spmv_csr_f64(int n_rows, int n_cols, int nnz, const int* rowptr, const int* col, const double* val,
                       const double* x, double* y, void* stream);
int impes_spmv_csr_f32(int n_rows, int n_cols, int nnz, const int* rowptr, const int* col, const float* val,
                       const float* x, float* y, void* stream);

/* y = A x for the 7-point pressure stencil in face form (diag d, face coefficients
 * ox/oy/oz stored at the lower cell of each face, z fastest). Replaces the sparse
 * product with the COO matrix of Simulator._build_pressure_matrix_vectorized,
 * src/simulator/simulation.py:1546-1580. */
int impes_spmv_st7_f64(int nx, int ny, int nz, const double* d, const double* ox, const double* oy,
                       const double* oz, const double* x, double* y, void* stream);

/* x_out = x_in + omega * inv_relax .* (b - A x_in): one fused L1-Jacobi sweep.
 * Replaces ClassicalAMG._smooth, src/solver/classical_amg.py:1016-1035. */
int impes_l1jacobi_csr_f64(int n, int nnz, const int* rowptr, const int* col, const double* val,
                           const double* inv_relax, double omega, const double* x_in, const double* b,
                           double* x_out, void* stream);
int impes_l1jacobi_csr_f32(int n, int nnz, const int* rowptr, const int* col, const float* val,
                           const float* inv_relax, float omega, const float* x_in, const float* b, float* x_out,
                           void* stream);

/* diag[i] = a_ii (0 if the row has no diagonal entry); used to form the row scaling
 * D^-1 of the Krylov solver. Replaces ClassicalAMG._extract_diag, classical_amg.py:940-956. */
int impes_csr_diagonal(int n, const int* rowptr, const int* col, const double* val, double* diag, void* stream);

/* out[0] = x . y (device scalar, FP64 accumulation, deterministic two-stage
 * reduction). Replaces the torch.dot / norm calls of the solvers,
 * src/solver/classical_amg.py:1136-1142, src/simulator/simulation.py:1677-1689. */
int impes_dot_f64(long long n, const double* x, const double* y, double* out, void* stream);

/* y = x (device, n doubles). No counterpart in the reference: the known-size
 * streaming kernel that calibrates HBM traffic counters. */
int impes_copy_f64(long long n, const double* x, double* y, void* stream);

/* ------------------------------------------------------------------ AMG hierarchy */

typedef struct {
  double theta;          /* strength threshold (reference default 0.25) */
  int max_levels;        /* reference max_levels (hierarchy holds up to max_levels+1 levels) */
  int coarsest_size;     /* stop coarsening at n <= coarsest_size */
  int pre_smooth;        /* V-cycle pre-sweeps (default 2; reference 3) */
  int post_smooth;       /* V-cycle post-sweeps (2) */
  double omega;          /* L1-Jacobi damping (default 0.8; reference 0.7) */
  double equil_threshold;/* equilibrate when the median |a_ii| < threshold (1e-6) */
  int dense_max;         /* coarsest level solved densely up to this size (500; at most 4096: above
                          * 160 rows the inverse is formed on the matrix cores, MFMA FP64, when
                          * the level is SPD; else by pivoted Gauss-Jordan up to 512 rows) */
  int coarse_sweeps;     /* Jacobi sweeps on a larger coarsest level (50) */
  int anchor;            /* forced C point on every level, -1 = none (reference anchor_idx) */
  /* Preconditioner choices that are not in the reference; they change the iteration
   * count, not the converged answer. The reference behaviour is equil_mode 0,
   * coarse_pin 1, trunc 0, pmax 0, coarsening 0, dominant 0. */
  int equil_mode;        /* 0 symmetric D^-1/2 A D^-1/2 (reference), 1 scalar A/median|a_ii| (default) */
  int coarse_pin;        /* coarsest DOF-0 pin: 0 never, 1 reference (multi-level), 2 only if singular (default) */
  double trunc;          /* drop interpolation weights below trunc * row max (default 0.2, 0 = off) */
  int pmax;              /* at most pmax interpolation weights per F row (default 4, 0 = unlimited) */
  int coarsening;        /* C/F split: 0 one pass + lift (reference), 1 iterated independent set (default) */
  double dominant;       /* rows with |a_ii| > dominant * sum|a_ij| become C points (default 10, 0 = off) */
} impes_amg_params;

void impes_amg_default_params(impes_amg_params* p);

/* Replaces ClassicalAMG.__init__ (hierarchy build), src/solver/classical_amg.py:699-938. */
int impes_amg_create(const impes_amg_params* params, impes_amg** out);
int impes_amg_destroy(impes_amg* amg);
int impes_amg_setup_csr(impes_amg* amg, int n, int nnz, const int* rowptr, const int* col, const double* val,
                        void* stream);
/* Same, from the 7-point stencil; level 0 then runs index-free. */
int impes_amg_setup_st7(impes_amg* amg, int nx, int ny, int nz, const double* d, const double* ox,
                        const double* oy, const double* oz, void* stream);
/* New values, same pattern. Replaces ClassicalAMG.update_matrix,
 * src/solver/classical_amg.py:1308-1392. */
int impes_amg_update_csr(impes_amg* amg, const double* val, void* stream);
int impes_amg_update_st7(impes_amg* amg, const double* d, const double* ox, const double* oy, const double* oz,
                         void* stream);

/* Introspection (host outputs): ClassicalAMG.levels, classical_amg.py:891-904. */
int impes_amg_num_levels(const impes_amg* amg, int* out);
int impes_amg_level_info(const impes_amg* amg, int level, int* n, int* nnz, int* p_nnz, int* is_coarsest,
                         int* dense);
int impes_amg_equilibrated(const impes_amg* amg, int* out);
/* The per-level V-cycle times of impes_model_amg_level_times for a hierarchy handle. */
int impes_amg_level_times(impes_amg* amg, int precision, int cycles, double* ms_host, double* bytes_host,
                          int max_levels, double* cycle_ms, void* stream);
int impes_amg_setup_seconds(const impes_amg* amg, double* out);
/* Copies level data into caller device buffers (sizes from level_info; any pointer may be NULL). */
int impes_amg_export_level(const impes_amg* amg, int level, int* rowptr, int* col, double* val, int* p_rowptr,
                           int* p_col, double* p_val, signed char* cf, double* inv_relax, void* stream);
/* dhalf_inv = D^-1/2 of the fine matrix when equilibrated, ones otherwise
 * (reference ClassicalAMG.Dhalf_inv, classical_amg.py:774-788). */
int impes_amg_export_scaling(const impes_amg* amg, double* dhalf_inv, void* stream);

/* One V(pre,post) cycle in the equilibrated space, x <- V(b); precision 64 or 32
 * (32: b and x are FP64 buffers, converted at the boundary).
 * Replaces ClassicalAMG._v_cycle, src/solver/classical_amg.py:1037-1106. */
int impes_amg_vcycle(impes_amg* amg, int precision, const double* b, double* x, int x_is_zero, void* stream);
/* The transposed cycle, V(post,pre): for the symmetric IMPES operator (R = P^T,
 * diagonal L1-Jacobi splitting, symmetric coarse solve) the zero-start cycle
 * satisfies V(pre,post)^T = V(post,pre). No counterpart in the reference; it is
 * the transposed preconditioner of an adjoint solve (the reference's V(3,2)
 * cycle is not self-adjoint; the default V(2,2) is). */
int impes_amg_vcycle_transposed(impes_amg* amg, int precision, const double* b, double* x, int x_is_zero,
                                void* stream);
/* Preconditioner of the row-scaled system D^-1 A: z = S V(S D v), S = D^-1/2 or I.
 * Replaces the apply closure of amg_preconditioner, src/solver/pressure_amg.py:155-158. */
int impes_amg_precondition(impes_amg* amg, int precision, const double* v, double* z, void* stream);
/* No counterpart in the reference: precondition() replays the V-cycle from a HIP graph
 * captured on first use after each setup (default; environment IMPES_GRAPHS=0 turns it
 * off for new hierarchies). enable 0 runs the same kernels as eager launches. */
int impes_amg_set_graphs(impes_amg* amg, int enable);

/* ------------------------------------------------------------------ FGMRES */

typedef struct {
  int iterations;
  int restarts;
  int converged;
  int precision;
  double rel_residual; /* ||D^-1 (b - A x)|| / ||D^-1 b|| at exit */
  double bnorm;        /* ||D^-1 b|| */
  double r0_rel;       /* relative residual of the initial guess */
  int checks;          /* pressure-error checks run (impes_solver_set_error_tol) */
  double err_estimate; /* last check's ||M r||_inf / ||x||_inf */
} impes_solve_stats;

int impes_solver_create(impes_solver** out);
int impes_solver_destroy(impes_solver* solver);
/* Pressure-error check of the FGMRES solves on this handle (0 = off, the default).
 * No counterpart in the reference, whose amg_solve stops on the residual alone
 * (classical_amg.py:1136-1142). Once the relative residual reaches rtol, one
 * preconditioner application to the scaled residual, z = M D^-1 (b - A x),
 * estimates the error A^-1 (b - A x); x += z refines the answer, and the solve
 * ends when ||z||_inf <= etol/2 * ||x||_inf. Otherwise the residual target is
 * lowered by the ratio of estimate to residual and the iteration continues; the
 * ratio is kept for the next solve on the handle, which starts from the
 * predicted target. */
int impes_solver_set_error_tol(impes_solver* solver, double etol);

/* FP64 restarted FGMRES on D^-1 A x = D^-1 b, right-preconditioned by one AMG
 * V-cycle (amg may be NULL: unpreconditioned). dinv = 1/diag(A). Takes over
 * amg_solve, src/solver/pressure_amg.py:35-89, whose stationary V-cycle
 * iteration it replaces by a Krylov method. stats is a host pointer. */
int impes_fgmres_csr(impes_solver* solver, impes_amg* amg, int precision, int n, int nnz, const int* rowptr,
                     const int* col, const double* val, const double* dinv, const double* b, double* x,
                     int use_x0, double rtol, int restart, int maxit, impes_solve_stats* stats, void* stream);
int impes_fgmres_st7(impes_solver* solver, impes_amg* amg, int precision, int nx, int ny, int nz,
                     const double* d, const double* ox, const double* oy, const double* oz, const double* dinv,
                     const double* b, double* x, int use_x0, double rtol, int restart, int maxit,
                     impes_solve_stats* stats, void* stream);

/* Stationary V-cycle iteration in physical space (x = D^-1/2-rescaled cycles of the
 * equilibrated matrix), stopping on ||bs - A xs||/||bs|| < tol after >= min_cycles.
 * Replaces ClassicalAMG.solve (classical_amg.py:1108-1155, use_x0 = x0 given,
 * min_cycles = 1) and ClassicalAMG.apply (classical_amg.py:1157-1180). rel and
 * cycles_done are host pointers (may be NULL). */
int impes_amg_cycles(impes_amg* amg, int precision, const double* b, double* x, int use_x0, int min_cycles,
                     double tol, int max_cycles, double* rel, int* cycles_done, void* stream);

/* The same iteration as impes_amg_cycles with no device allocation: its workspace
 * (three FP64 n-vectors and the reduction partials) is a static device array of the
 * library, 24 MiB per GPU, held from library load to process exit. No counterpart in
 * the reference; it is for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc). Serves hierarchies of up to 1,048,576 rows and
 * fails with an error above that. All calls on one GPU share the workspace: calls
 * that overlap on one GPU (other threads, other streams) race. Nothing in the
 * workspace carries over from one call to the next. */
int impes_amg_cycles_noalloc(impes_amg* amg, int precision, const double* b, double* x, int use_x0,
                             int min_cycles, double tol, int max_cycles, double* rel, int* cycles_done,
                             void* stream);

/* ------------------------------------------------------------------ IMPES model */

typedef struct {
  int nx, ny, nz;     /* Reservoir.dimensions, reservoir.py:18-19 */
  double dx, dy, dz;  /* Reservoir.grid_size, reservoir.py:22-23 */
} impes_grid;

/* Fluid.__init__ parameters (fluid.py:26-110) and scheme knobs (simulation.py:826-828,1058,1354,1377). */
typedef struct {
  int stone2;                          /* relative_permeability.model == "stone2" */
  double nw, no, ng, ko_end_w, ko_end_g, now_, nog;
  double sw_cr, so_r, sg_cr;
  double mu_o, mu_w, mu_g;             /* Pa s, used without a PVT table */
  double rho_o_ref, rho_w_ref, rho_g_ref;
  double c_o, c_w, c_g, c_r, cf;       /* 1/Pa; cf = (c_o+c_w+c_g+c_r)/3 */
  double pc_scale, pc_exp, pcog_scale, pcog_exp;
  double g;                            /* 9.81 */
  double max_ds;                       /* max_saturation_change */
  int max_substeps;
  double rate_scale;                   /* global_rate_scale */
} impes_fluid;

/* One well (well.py:10-55); cell is the z-fastest flat index (simulation.py:816). */
typedef struct {
  int cell;
  int producer;       /* 0 injector, 1 producer */
  int control;        /* 0 rate (m3/day), 1 bhp (MPa), 2 ignored */
  int gas;            /* injected phase: 0 water, 1 gas */
  int surface;        /* rate_type == "surface" */
  int surface_phase;  /* 0 oil, 1 water, 2 gas, 3 liquid */
  int has_bhp_min;
  int lim_mask;       /* bit0 wopr, bit1 wlpr, bit2 wgpr, bit3 liqr */
  double value;
  double wi;          /* Peaceman index of well.py:44-51 */
  double bhp_min;     /* MPa */
  double lim[4];      /* m3/day */
} impes_well;

/* Creates the device model of one reservoir. pvt_host (host) holds 9*pvt_n doubles:
 * pressure_MPa, Bo, Bw, Bg, mu_o_cP, mu_w_cP, mu_g_cP, Rs, Rv (pvt.py:18-46); pvt_n = 0: no table.
 * Takes over Simulator.__init__ (simulation.py:19-106) for the IMPES branch. */
int impes_model_create(const impes_grid* grid, const impes_fluid* fluid, int pvt_n, const double* pvt_host,
                       int nwells, const impes_well* wells_host, impes_model** out);
int impes_model_destroy(impes_model* m);
/* kx, ky, kz in m^2 and pore volume in m^3 per cell (device): transmissibilities of
 * simulation.py:52-74 computed on the device. */
int impes_model_set_rock(impes_model* m, const double* kx, const double* ky, const double* kz, const double* pv,
                         void* stream);
/* New well set (ReservoirEnv._apply_action changes controls every step, rl_env.py:129-140):
 * uploaded on `stream` (the host array may be released on return). */
int impes_model_set_wells(impes_model* m, int nwells, const impes_well* wells_host, void* stream);
/* Batched environments (the reference steps one reservoir per Simulator inside
 * ReservoirEnv.step, rl_env.py:90-108): the grid holds nx / env_nx independent
 * reservoirs of env_nx x-layers each, stacked along x. Faces between them carry no
 * flow and every reservoir gets its own substep count (the saturation report then
 * ends with those nx / env_nx counts). env_nx 0 or nx: one reservoir. Call before
 * impes_model_set_rock. */
int impes_model_set_envs(impes_model* m, int env_nx, void* stream);
int impes_model_set_amg(impes_model* m, const impes_amg_params* params);
int impes_model_report_len(const impes_model* m, int* out);
/* Pressure equation at (p, sw, sg): Simulator._impes_pressure_step up to the solve
 * (simulation.py:867-902): mobilities, upwinding, _calculate_well_terms,
 * _build_pressure_matrix_vectorized, _build_pressure_rhs. */
int impes_model_assemble(impes_model* m, double dt, const double* p, const double* sw, const double* sg,
                         void* stream);
/* AMG-preconditioned FP64 FGMRES on the assembled system (AMG setup on the first call,
 * numeric update after); p_prev is the initial guess, p_new the solution.
 * Replaces the "amg" branch of simulation.py:904-929. */
int impes_model_solve_pressure(impes_model* m, int precision, double rtol, int restart, int maxit,
                               const double* p_prev, double* p_new, impes_solve_stats* stats, void* stream);
/* The pressure-error check of impes_solver_set_error_tol for this model's solves
 * (0 = off, the default): holds the pressure to etol relative to the exact solve
 * where the residual stop alone would not. No counterpart in the reference. */
int impes_model_set_error_tol(impes_model* m, double etol);
/* Jacobi-preconditioned CG, reference semantics (absolute stop on sqrt(r.z), x0 = 0):
 * Simulator._solve_pressure_cg_pytorch, simulation.py:1662-1697. converged/iterations: host. */
int impes_model_solve_cg(impes_model* m, double tol, int maxit, double* p_new, int* converged, int* iterations,
                         void* stream);
/* Explicit transport: Simulator._impes_saturation_step (simulation.py:968-1544).
 * report_host receives impes_model_report_len doubles (substeps, ledgers, per-well diagnostics). */
int impes_model_saturation(impes_model* m, double dt, const double* p_prev, const double* p_new, const double* sw,
                           const double* sg, double* sw_out, double* sg_out, double* so_out, double* report_host,
                           void* stream);
/* ---- reverse sweep of the step (the hand-written adjoint). No counterpart in the
 * reference (its ROADMAP.md:53 lists an adjoint as future work). One reverse step n,
 * for the step from (p_prev, sw, sg) to (p_new, sw_next, sg_next):
 *   impes_model_assemble(m, dt, p_prev, sw, sg)       re-assemble step n;
 *   impes_model_transport_vjp(...)                    transport reversed;
 *   impes_model_solve_adjoint(m, ..., g_p_new, lam)   lam = A^-1 g_p_new (A = A^T);
 *   impes_model_assemble_vjp(..., lam, ...)           assembly reversed.
 * Cotangents of the face transmissibilities and well controls accumulate in the
 * model from impes_model_adjoint_reset on; impes_model_rock_vjp turns the former into
 * permeability cotangents, impes_model_adjoint_values returns the latter.
 * Branches (upwinding, substep count, clamps, chokes, limits, bhp_min) are those of
 * the forward values. Batched environments are not supported. */
int impes_model_adjoint_reset(impes_model* m, void* stream);
/* Replaces the transport part of the reverse sweep that the reference lacks for
 * Simulator._impes_saturation_step (simulation.py:968-1544). Inputs: g_sw_next,
 * g_sg_next (cotangents of the step's output saturations), g_obs_host (host,
 * nwells x 3: cotangents of the producers' surface oil/water/gas rates, the
 * report's observation slots). Writes g_sw, g_sg (cotangents of sw, sg); adds into
 * g_p_new (cotangent of p_new) and g_p_prev. Synchronises the stream. */
int impes_model_transport_vjp(impes_model* m, double dt, const double* p_prev, const double* p_new,
                              const double* sw, const double* sg, const double* g_sw_next, const double* g_sg_next,
                              const double* g_obs_host, double* g_p_new, double* g_p_prev, double* g_sw, double* g_sg,
                              void* stream);
/* Transposed pressure solve of the assembled step (the operator is symmetric: the
 * forward's hierarchy, numerically updated, inside FP64 FGMRES from a zero start). */
int impes_model_solve_adjoint(impes_model* m, int precision, double rtol, int restart, int maxit, const double* g,
                              double* lam, impes_solve_stats* stats, void* stream);
/* Reverse of the pressure assembly (simulation.py:867-902, 1546-1660) given the
 * adjoint solution lam and the solved p_new: adds into g_p_prev, g_sw, g_sg. */
int impes_model_assemble_vjp(impes_model* m, double dt, const double* p_prev, const double* sw, const double* sg,
                             const double* p_new, const double* lam, double* g_p_prev, double* g_sw, double* g_sg,
                             void* stream);
/* Reverse of the transmissibilities (simulation.py:52-74): accumulated face
 * cotangents -> cotangents of kx, ky, kz (m^2) per cell. */
int impes_model_rock_vjp(impes_model* m, const double* kx, const double* ky, const double* kz, double* g_kx,
                         double* g_ky, double* g_kz, void* stream);
/* Accumulated cotangents of the well control values (host, nwells). */
int impes_model_adjoint_values(impes_model* m, double* g_values_host, void* stream);
/* Copies a model field into out (device, n cells): 0 diag, 1 ox, 2 oy, 3 oz, 4 rhs, 5 accumulation,
 * 6 gravity+capillary rhs, 7 tx, 8 ty, 9 tz, 10 pore volume, 11 total mobility, 12 1/diag. */
int impes_model_export(impes_model* m, int field, double* out, void* stream);
/* Elementwise fluid properties (fluid.py:129-248,273-296,330-435): out is 12 x n:
 * kro krw krg pc_ow pc_og ct rho_w rho_o rho_g mu_o mu_w mu_g (Pa s). */
int impes_model_eval_props(impes_model* m, long long n, const double* p, const double* sw, const double* sg,
                           double* out, void* stream);
/* PVT lookup (pvt.py:52-64 / fluid.py:182-188), key 0..7 = Bo Bw Bg mu_o mu_w mu_g Rs Rv. */
int impes_model_eval_pvt(impes_model* m, int key, long long n, const double* p, double* out, void* stream);
/* Hierarchy summary of the pressure solver (after the first solve). */
int impes_model_amg_levels(const impes_model* m, int* num_levels, int* sizes_host, int* nnz_host, int max_levels);
/* Wall time of the last hierarchy build of the pressure solver (0 before the first
 * solve): the setup part of ClassicalAMG.__init__ (classical_amg.py:699-938), which
 * the first pressure solve of a run performs and later solves update numerically. */
int impes_model_amg_setup_seconds(const impes_model* m, double* out);
/* Measurement, no counterpart in the reference: `cycles` eager V-cycles of the
 * pressure solver's hierarchy on its current V-cycle input, with HIP events between
 * the phases of every level. ms_host[l] (l < min(levels, max_levels)) = mean stream
 * time per cycle of level l's own launches (sweeps, residual, restriction,
 * prolongation, or the coarsest solve; children excluded); bytes_host[l] (may be
 * NULL) = the algorithmic HBM bytes of those launches per visit; *cycle_ms = mean
 * time of the same cycle replayed from its HIP graph. precision 32 or 64. */
int impes_model_amg_level_times(impes_model* m, int precision, int cycles, double* ms_host, double* bytes_host,
                                int max_levels, double* cycle_ms, void* stream);

/* ------------------------------------------------------------------ measurement */
/* No counterpart in the reference (instrumentation). Times every launch of one
 * kernel, by name ("st7_jacobi_f32", "st7_jacobi_f64", "csr_jacobi_f32",
 * "csr_jacobi_f64", "csr_residual_f32", "st7_residual_f32",
 * "st7_rowscaled_spmv_f64"), with a HIP event pair on its launching stream, and
 * adds the algorithmic bytes of each launch. config resets the totals; kernel
 * NULL or enable 0 switches timing off. read waits for the recorded events. */
int impes_ktimer_config(const char* kernel, int enable);
int impes_ktimer_read(double* ms, double* bytes, long long* launches);

#ifdef __cplusplus
}
#endif

#endif /* IMPES_HIP_H */
