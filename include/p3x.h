/*
 * p3x.h -- C ABI of the MI355X P3 microphysics step (drop-in for EAMxx p3_main).
 *
 * Reference interface replaced (paths under components/eamxx/src/physics/p3/):
 *   Functions<Real,Device>::p3_main(runtime, prog, diag_in, diag_out, infra,
 *       hist, tables, workspace_mgr, nj, nk)            p3_functions.hpp:1356-1369
 *   P3Runtime                                           p3_functions.hpp:112-158
 *   P3PrognosticState                                   p3_functions.hpp:161-183
 *   P3DiagnosticInputs                                  p3_functions.hpp:186-214
 *   P3DiagnosticOutputs                                 p3_functions.hpp:217-241
 *   P3Infrastructure                                    p3_functions.hpp:244-264
 *   P3HistoryOnly                                       p3_functions.hpp:268-276
 *   P3LookupTables                                      p3_functions.hpp:279-290
 *   init_kokkos_ice_lookup_tables / init_kokkos_tables  impl/p3_table_ice_impl.hpp:16-78,
 *                                                       impl/p3_table3_impl.hpp:85-148
 *
 * Data layout.  Every per-level field is a DEVICE pointer, k = 0 at the model
 * top, 0 <= i < nj, in one of two layouts (p3x_launch.layout):
 *   P3X_COLUMN_MAJOR  the reference's Kokkos views (column, pack) LayoutRight
 *                     (eamxx_p3_process_interface.cpp:254-265): element
 *                     (column i, level k) at [i*ld + k], ld >= nk; the two
 *                     precipitation-flux fields have nk+1 levels (interfaces)
 *                     at [i*ld_iface + k], ld_iface >= nk+1.  A pack-padded
 *                     view passes its padded extents as ld and ld_iface.
 *   P3X_LEVEL_MAJOR   (the default) element (i, k) at [k*ld + i], ld >= nj,
 *                     flux rows k = 0..nk at the same stride.
 * precip_liq_surf / precip_ice_surf hold nj values.  Both layouts give the
 * same results bit for bit; the kernels map their lanes so that loads
 * coalesce in either.
 *
 * Field names, units and meanings are those of the reference structs.
 * Semantics follow the reference's p3_main exactly, with one documented
 * choice: all nk+1 entries of precip_liq_flux / precip_ice_flux are zeroed
 * (as the Fortran reference does); see p3x_infrastructure for the exchange
 * fields.
 *
 * Thread safety: one context may be used by one host thread at a time.
 * Every call is enqueued on the stream given in p3x_launch (default stream if
 * NULL) and, unless p3x_launch.async is set, is complete on return.  The
 * p3x_main calls of one context share its scratch: a call on another stream
 * than the previous call's is made to wait on the device for that call, so
 * asynchronous calls on several streams run one after another.
 */
#ifndef P3X_H
#define P3X_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
enum {
  P3X_OK = 0,
  P3X_ERR_ARG = 1,          /* null pointer, bad size or bad option                 */
  P3X_ERR_TABLE_FILE = 2,   /* cannot open / parse the ice lookup table             */
  P3X_ERR_TABLE_VERSION = 3,/* table header is not "VERSION 4.1.1"                 */
  P3X_ERR_DEVICE = 4,       /* a HIP runtime call failed                           */
  P3X_ERR_TEMPERATURE = 5,  /* T <= 0 or NaN reached the saturation functions;     */
                            /* the reference aborts here (physics_saturation_impl.hpp:14-36) */
  P3X_ERR_NO_MEMORY = 6,    /* device allocation failed                            */
  P3X_ERR_SUBSTEPS = 7,     /* a sedimentation loop hit the 100000-substep cap     */
  P3X_ERR_BUDGET = 8,       /* served form: a member does not fit the HBM budget   */
  P3X_ERR_CHECK_VALUES = 9, /* with p3x_launch.check_values: a level's T left     */
                            /* (173, 323) K after the step (p3_check_values_impl.hpp:25-79) */
  P3X_ERR_SCRATCH_LIMIT = 10 /* p3x_main_noalloc: the column set needs more scratch */
                            /* than its fixed array holds (P3X_NOALLOC_SCRATCH_BYTES) */
};

/* ---- P3Runtime (p3_functions.hpp:112-158); defaults from p3x_runtime_defaults ---- */
typedef struct p3x_runtime {
  double max_total_ni;
  double autoconversion_prefactor;
  double autoconversion_qc_exponent;
  double autoconversion_nc_exponent;
  double autoconversion_radius;
  double accretion_prefactor;
  double accretion_qc_exponent;
  double accretion_qr_exponent;
  double rain_selfcollection_prefactor;
  double rain_selfcollection_breakup_diameter;
  double constant_mu_rain;
  double spa_ccn_to_nc_factor;
  double cldliq_to_ice_collection_factor;
  double rain_to_ice_collection_factor;
  double min_rime_rho;
  double max_rime_rho;
  double immersion_freezing_exponent;
  double deposition_nucleation_exponent;
  double ice_sedimentation_factor;
  int do_ice_production;
} p3x_runtime;

/* ---- P3PrognosticState (p3_functions.hpp:161-183), in/out ---- */
typedef struct p3x_prognostic_state {
  double *qc, *nc, *qr, *nr, *qi, *qm, *ni, *bm, *qv, *th;
} p3x_prognostic_state;

/* ---- P3DiagnosticInputs (p3_functions.hpp:186-214) ---- */
typedef struct p3x_diagnostic_inputs {
  const double *nc_nuceat_tend, *nccn, *ni_activated, *inv_qc_relvar;
  const double *cld_frac_i, *cld_frac_l, *cld_frac_r;
  const double *pres, *dz, *dpres, *inv_exner, *qv_prev, *t_prev;
} p3x_diagnostic_inputs;

/* ---- P3DiagnosticOutputs (p3_functions.hpp:217-241) ---- */
typedef struct p3x_diagnostic_outputs {
  double *qv2qi_depos_tend;
  double *precip_liq_surf, *precip_ice_surf;             /* nj values          */
  double *diag_eff_radius_qc, *diag_eff_radius_qi, *diag_eff_radius_qr;
  double *rho_qi;
  double *precip_liq_flux, *precip_ice_flux;             /* nk+1 rows          */
  double *precip_total_tend, *nevapr;
} p3x_diagnostic_outputs;

/* ---- P3Infrastructure (p3_functions.hpp:244-264) ----
 * its/ite/kts/kte and col_location are carried for interface parity; the step
 * uses nj/nk of p3x_main.  The reference's p3_main leaves the three exchange
 * fields of P3HistoryOnly untouched where a column or level is skipped and
 * accumulates into vap_liq_exchange in part 3, so its caller
 * (P3Microphysics::run_impl, eamxx_p3_run.cpp:27-29) zeroes them first.
 * exchange_zeroed_on_entry = 1 declares that the caller would have zeroed them:
 * the kernel then writes every entry and never reads them (fewer bytes). */
typedef struct p3x_infrastructure {
  double dt;
  int it, its, ite, kts, kte;
  int predictNc;
  int prescribedCCN;
  const double* col_location;
  int exchange_zeroed_on_entry;
} p3x_infrastructure;

/* ---- P3HistoryOnly (p3_functions.hpp:268-276) ---- */
typedef struct p3x_history_only {
  double *liq_ice_exchange, *vap_liq_exchange, *vap_ice_exchange;
} p3x_history_only;

/* ---- P3LookupTables contents, HOST arrays, row-major (p3_functions.hpp:279-290) ----
 *   ice_table     [5][4][50][12]   collect_table [5][4][50][30][2] (log10 values)
 *   vn/vm/revap   [300][10]        mu_r_table    [150]        dnu [16]            */
typedef struct p3x_table_data {
  double ice_table[5 * 4 * 50 * 12];
  double collect_table[5 * 4 * 50 * 30 * 2];
  double vn_table[300 * 10];
  double vm_table[300 * 10];
  double revap_table[300 * 10];
  double mu_r_table[150];
  double dnu[16];
} p3x_table_data;

/* ---- launch options (no reference counterpart: how the step runs on MI355X) ---- */
enum { P3X_FP64 = 0, P3X_MIXED = 1 };
enum { P3X_LEVEL_MAJOR = 0, P3X_COLUMN_MAJOR = 1 };
typedef struct p3x_launch {
  int precision;           /* P3X_FP64 (drop-in) or P3X_MIXED (FP32 bulk rates)  */
  int ld;                  /* level-major: row stride (>= nj; 0 means nj);        */
                           /* column-major: column stride (>= nk; 0 means nk)     */
  /* part 1 kernel: columns per block x threads per block, both 0 for the
   * default, else one of 16x128, 16x256, 32x256, 64x256, 32x512, 64x512      */
  int columns_per_block;
  int threads_per_block;
  /* sedimentation kernels: lanes per column, 0 for each species group's
   * default, else 16, 32 or 64 for both groups; nk may not exceed 8 x this
   * (nk <= 512 in any case)                                                   */
  int sed_lanes_per_column;
  /* part 2 kernel: wavefronts per SIMD its registers must allow, 0 for the
   * default (2), else 1 or 2; an explicit block shape without it means 1    */
  int part12_waves_per_simd;
  void* stream;            /* hipStream_t; NULL = default stream                  */
  int async;               /* 1: return after enqueueing (status from p3x_sync)   */
  /* optional device arrays for the integer diagnostics (NULL = off):
   *   col_status [nj]        0 skipped after part 1, 1 after part 2, 2 full
   *   sed_substeps [3*nj]    CFL substeps of cloud, rain, ice sedimentation
   *   ice_table_idx          part-3 ice-table cell jj*10000+ii*100+i or -1, a
   *                          level field in the layout of the others             */
  int* col_status;
  int* sed_substeps;
  int* ice_table_idx;
  int layout;              /* P3X_LEVEL_MAJOR (0) or P3X_COLUMN_MAJOR              */
  int ld_iface;            /* column-major: column stride of the two nk+1-level   */
                           /* flux fields (0 means max(ld, nk+1)); level-major: 0 */
  /* optional device array [nj] (NULL = off): the reference's debug-build
   * check_values after part 3 (p3_main_impl.hpp:316-324, bounds at
   * p3_check_values_impl.hpp:25-79) on the fully stepped columns: bit 0 = some
   * level's T = th exner is outside (173, 323) K (the trap: the step then
   * returns P3X_ERR_CHECK_VALUES), bit 1 = some level's qv is outside
   * (0, 40e-3) (a flag only, as in the reference)                             */
  int* check_values;
} p3x_launch;

typedef struct p3x_context p3x_context;

/* Fill rt with the P3Runtime defaults (p3_functions.hpp:114-133). */
void p3x_runtime_defaults(p3x_runtime* rt);
/* Fill launch options with defaults (FP64, tuned block shape, synchronous). */
void p3x_launch_defaults(p3x_launch* opts);

/* Host: read the ice lookup table text (reader of p3_table_ice_impl.hpp:16-78,
 * including the "VERSION 4.1.1" check) and build the rain tables (the
 * reference generates them with micro_p3.F90 p3_init_b, :218-341) and dnu
 * (p3_table3_impl.hpp:120-135). */
int p3x_tables_load(const char* ice_table_file, p3x_table_data* out);

/* Create a context on the current HIP device: uploads the tables. */
int p3x_init(const p3x_table_data* tables, p3x_context** ctx);
void p3x_finalize(p3x_context* ctx);
/* Allocate the context's scratch for column sets up to ld columns (the
 * leading dimension) by nk levels, as initialize_impl sizes the process's
 * buffers once (eamxx_p3_process_interface.cpp:124-217, the
 * requested_buffer_size_in_bytes / init_buffers pair): after it, p3x_main
 * allocates nothing for shapes within it.  Without it, p3x_main allocates on
 * the first call and whenever a larger shape arrives. */
int p3x_reserve(p3x_context* ctx, int ld, int nk);

/* One P3 step over nj columns x nk levels (p3_functions.hpp:1356-1369).
 * elapsed_us (optional) receives the device time of the step (HIP events),
 * the counterpart of p3_main's returned microseconds (p3_main_impl.hpp:328-330). */
int p3x_main(p3x_context* ctx, const p3x_runtime* runtime,
             const p3x_prognostic_state* prognostic_state,
             const p3x_diagnostic_inputs* diagnostic_inputs,
             const p3x_diagnostic_outputs* diagnostic_outputs,
             const p3x_infrastructure* infrastructure,
             const p3x_history_only* history_only,
             int nj, int nk, const p3x_launch* opts, double* elapsed_us);

/* No counterpart in the reference: p3x_main for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, kernels and results as p3x_main;
 * the scratch of the step (7 x nk x nj doubles + nj ints, rounded up to 256
 * bytes) comes from a fixed device array of P3X_NOALLOC_SCRATCH_BYTES that the
 * library holds on every GPU it runs on until the process ends, instead of
 * from the context, and the call allocates nothing (no hipMalloc, hipFree or
 * allocating library routine).  A column set needing more returns
 * P3X_ERR_SCRATCH_LIMIT without launching (21600 x 72 and 16384 x 128 fit).
 * Every p3x_main_noalloc call on one GPU, from any context or stream, shares
 * that array, which keeps the last call's contents: calls that overlap on one
 * GPU race (order them on one stream, or with events). */
#define P3X_NOALLOC_SCRATCH_BYTES 134217728ull /* 128 MiB */
int p3x_main_noalloc(p3x_context* ctx, const p3x_runtime* runtime,
                     const p3x_prognostic_state* prognostic_state,
                     const p3x_diagnostic_inputs* diagnostic_inputs,
                     const p3x_diagnostic_outputs* diagnostic_outputs,
                     const p3x_infrastructure* infrastructure,
                     const p3x_history_only* history_only,
                     int nj, int nk, const p3x_launch* opts, double* elapsed_us);

/* Wait for the context's enqueued work and return its status. */
int p3x_sync(p3x_context* ctx);

/* Device time of the context's last p3x_main (waits for it) in two intervals:
 * the part 1 and part 2 kernels, then the sedimentation and part 3 kernels.  No reference counterpart (the
 * reference times the whole p3_main, p3_main_impl.hpp:328-330); it splits that
 * figure for profiling. */
int p3x_last_kernel_times(p3x_context* ctx, double* part12_us, double* sed_us);

/* ---- batched form (no reference counterpart: the reference runs one column
 * set per p3_main call).  Many independent column sets -- each with its own
 * nj, nk, dt and flags -- advance one P3 step each in one launch per kernel,
 * optionally captured once into a HIP graph and replayed.  Every set's
 * results equal those of a p3x_main call on that set alone (same kernels,
 * same per-column arithmetic). ---- */
typedef struct p3x_column_set {
  p3x_prognostic_state prognostic_state;
  p3x_diagnostic_inputs diagnostic_inputs;
  p3x_diagnostic_outputs diagnostic_outputs;
  p3x_history_only history_only;
  p3x_infrastructure infrastructure;  /* dt, predictNc, prescribedCCN, exchange_zeroed_on_entry */
  int nj, nk, ld;                     /* ld as in p3x_launch (0 = dense)                     */
  int *col_status, *sed_substeps, *ice_table_idx;  /* optional, as in p3x_launch         */
  int layout, ld_iface;               /* as in p3x_launch                                    */
  int* check_values;                  /* optional, as in p3x_launch                          */
} p3x_column_set;

typedef struct p3x_batch_options {
  int precision;             /* P3X_FP64 or P3X_MIXED                                     */
  int columns_per_block;     /* part 1 shape, as in p3x_launch (0 = default)              */
  int threads_per_block;
  int sed_lanes_per_column;  /* 0 = default                                               */
  int part12_waves_per_simd; /* 0 = default, 1 or 2, as in p3x_launch                    */
  int blocks_per_cu;         /* persistent grid of each launch: this many blocks per CU;  */
                             /* 0 = as many as fit (occupancy); -1 = one block per unit   */
                             /* (the default: the kernels then run without their loop)   */
  int use_graph;             /* 1 = capture the launches into a HIP graph once, replay it */
} p3x_batch_options;

typedef struct p3x_batch p3x_batch;

void p3x_batch_defaults(p3x_batch_options* opts);
/* The descriptors are copied; the arrays they point to must stay valid while
 * the batch exists.  Allocates the batch's scratch and (use_graph) its graph. */
int p3x_batch_create(p3x_context* ctx, const p3x_runtime* runtime, const p3x_column_set* sets,
                     int nsets, const p3x_batch_options* opts, p3x_batch** batch);
/* Enqueue one step of every set on stream (NULL = default); returns after
 * enqueueing.  p3x_sync(ctx) waits and returns the status. */
int p3x_batch_run(p3x_batch* batch, void* stream);
/* Autotuner: times candidate shapes (part 1 + part 2 block shape, sedimentation
 * lanes per column, persistent blocks per CU) on the batch's own data, keeps the
 * fastest, and leaves every set's state as it was.  reps timed steps per
 * candidate; chosen / best_us (optional) receive the result. */
int p3x_batch_tune(p3x_batch* batch, void* stream, int reps, p3x_batch_options* chosen,
                   double* best_us);
void p3x_batch_destroy(p3x_batch* batch);

/* ---- the P3Microphysics atmosphere process (eamxx_p3_process_interface.hpp:22-444,
 * eamxx_p3_run.cpp:5-45): pre-amble, p3_main, post-amble on the field-manager
 * fields.  Names are the field names the process requests
 * (eamxx_p3_process_interface.cpp:24-121); every per-level array is a device
 * array in the layout the process was created with, per-column fields have ncol
 * entries. ---- */
typedef struct p3x_atm_fields {
  /* required inputs */
  const double *p_mid, *p_dry_mid, *pseudo_density, *pseudo_density_dry, *cldfrac_tot;
  const double *nc_nuceat_tend, *nccn, *ni_activated, *inv_qc_relvar;  /* nccn: prescribedCCN */
  /* updated (wet mixing ratios in and out) */
  double *T_mid, *qv, *qc, *nc, *qr, *nr, *qi, *qm, *ni, *bm;
  double *qv_prev_micro_step, *T_prev_micro_step;
  double *precip_liq_surf_mass, *precip_ice_surf_mass;             /* ncol, accumulated */
  /* computed */
  double *eff_radius_qc, *eff_radius_qi, *eff_radius_qr;           /* microns          */
  double *precip_total_tend, *nevapr, *rainfrac;
  double *micro_liq_ice_exchange, *micro_vap_liq_exchange, *micro_vap_ice_exchange;
  /* boundary fluxes of the column conservation checks (ncol each); NULL = off */
  double *vapor_flux, *water_flux, *ice_flux, *heat_flux;
} p3x_atm_fields;

typedef struct p3x_process p3x_process;

/* set_grids + initialize_impl: one process for ncol x nlev columns; allocates
 * the process-local arrays (inv_exner, th_atm, cld_frac_l/i, dz, the
 * precipitation fluxes, qv2qi_depos_tend, rho_qi) in the fields' layout.
 * p3x_process_create_layout takes the layout of the field-manager fields:
 * P3X_COLUMN_MAJOR with the views' column strides (ld for the nlev-level
 * fields, ld_iface for nlev+1-level ones; 0 = dense), as get_view<Pack**>()
 * hands them to the reference (eamxx_p3_process_interface.cpp:254-265);
 * p3x_process_create is the level-major, dense form. */
int p3x_process_create(p3x_context* ctx, const p3x_runtime* runtime, int ncol, int nlev,
                       int predictNc, int prescribedCCN, p3x_process** process);
int p3x_process_create_layout(p3x_context* ctx, const p3x_runtime* runtime, int ncol, int nlev,
                              int predictNc, int prescribedCCN, int layout, int ld, int ld_iface,
                              p3x_process** process);
/* run_impl(dt): pre-amble, p3_main, post-amble (same stream, same call); the
 * pre-amble runs inside p3_main's first kernel and the post-amble inside its
 * last, so a step launches the same kernels as p3x_main.  The layout is the
 * process's own: opts->layout, ld and ld_iface must be 0 or equal to it. */
int p3x_process_run(p3x_process* process, const p3x_atm_fields* fields, double dt,
                    const p3x_launch* opts);
void p3x_process_destroy(p3x_process* process);

/* ---- the mass-and-energy column conservation check that runs after the process
 * (share/property_checks/mass_and_energy_column_conservation_check.cpp:110-353).
 * Per-level device arrays in the given layout and stride (as p3x_launch; ld 0 =
 * dense); levels are summed k = 0 .. nlev-1.  Both calls are enqueued on
 * stream (NULL = default) and return at once. ---- */
typedef struct p3x_budget_fields {
  const double *pseudo_density, *T_mid, *u, *v;  /* u, v: horiz_winds; NULL = 0     */
  const double *qv, *qc, *qi, *qr;
  const double *ps, *phis;                       /* ncol each; NULL = 0             */
  int layout, ld;                                /* P3X_LEVEL_MAJOR (0) or column-major */
} p3x_budget_fields;

/* Column total mass [kg/m2] and total energy [J/m2] (:283-301, :312-340). */
int p3x_column_budget(p3x_context* ctx, const p3x_budget_fields* fields, int ncol, int nlev,
                      double* mass, double* energy, void* stream);
/* Per-column relative errors of a step against its boundary fluxes over dt
 * (:151-154, :171-174, fluxes :303-310, :341-353). */
int p3x_conservation_errors(p3x_context* ctx, const double* mass_before, const double* energy_before,
                            const double* mass_after, const double* energy_after,
                            const double* vapor_flux, const double* water_flux, const double* ice_flux,
                            const double* heat_flux, int ncol, double dt, double* mass_err,
                            double* energy_err, void* stream);

/* ---- the CldFraction process that runs before P3 and produces its cldfrac_tot
 * (physics/cld_fraction/eamxx_cld_fraction_process_interface.cpp:77-90,
 * cld_fraction_main_impl.hpp:12-95): ice cloud fraction 1 where qi exceeds a
 * threshold (strictly) and 0 elsewhere, total = max(ice, liquid), for the
 * model's threshold (default 1e-12) and the analysis threshold (default 1e-5).
 * Pointwise over ncol*nlev dense device arrays, so either layout (all six
 * arrays in the same one); enqueued on stream (NULL = default), returns at
 * once. ---- */
int p3x_cld_fraction(p3x_context* ctx, const double* qi, const double* cldfrac_liq, int ncol, int nlev,
                     double ice_threshold, double ice_4out_threshold, double* cldfrac_ice,
                     double* cldfrac_tot, double* cldfrac_ice_4out, double* cldfrac_tot_4out,
                     void* stream);

/* ---- SHOC's post-process, the producer of P3's inv_qc_relvar
 * (physics/shoc/eamxx_shoc_process_interface.hpp:272-300): inv_qc_relvar =
 * min(10, max(0.001, qc^2 / qc2)) where qc != 0 and qc2 != 0, else 1 (qc: the
 * cloud liquid mixing ratio, qc2: SHOC's variance of it; the wet/dry factor
 * cancels in the ratio), and cldfrac_liq clipped to at most 1 in place (NULL:
 * left alone).  Pointwise over ncol*nlev dense device arrays in one layout;
 * enqueued on stream (NULL = default), returns at once. ---- */
int p3x_shoc_postprocess(p3x_context* ctx, const double* qc, const double* qc2, double* cldfrac_liq,
                         double* inv_qc_relvar, int ncol, int nlev, void* stream);

/* ---- served form (no reference counterpart): step requests from concurrent
 * ensemble members, continuously batched.  Members register at any time, also
 * while others are being served; their arrays (caller-owned device memory) stay
 * resident and are advanced in place.  A member is admitted only while the
 * bytes all members' arrays span plus the slots' scratch fit the HBM budget.  A
 * scheduler thread admits queued requests into batches as launch slots and
 * scratch free up, one request per member in flight at a time, in arrival
 * order otherwise; batches on different slots run concurrently on their own
 * HIP streams.  Each request has its own status (a step that fails -- T <= 0
 * or NaN -- fails alone).  A request's result equals that of a p3x_main call
 * on its member alone. ---- */
typedef struct p3x_server_options {
  int precision;              /* P3X_FP64 or P3X_MIXED                                */
  int slots;                  /* batches in flight at once (one stream each); 0 = 4   */
  int max_batch;              /* requests per batch; 0 = 32                           */
  double scratch_budget_bytes;/* device scratch of all slots; 0 = 8 GiB               */
  int columns_per_block, threads_per_block, sed_lanes_per_column;  /* 0 = default  */
  int part12_waves_per_simd;                                        /* 0 = default  */
  double hbm_budget_bytes;    /* members' resident arrays + scratch; 0 = 200e9       */
} p3x_server_options;

typedef struct p3x_completion {
  unsigned long long tag;     /* as given to p3x_server_submit                        */
  int member;
  int status;                 /* P3X_OK or the step's error status                    */
  double latency_us;          /* submit -> completion seen by the scheduler           */
} p3x_completion;

typedef struct p3x_server_stats {
  long long submitted, completed;
  double p50_us, p99_us, max_us;  /* over all completed requests                      */
  double column_levels;           /* total column-levels stepped                      */
  double busy_s;                  /* first submit -> last completion                  */
  double resident_bytes;          /* bytes the registered members' arrays span        */
  double scratch_bytes;           /* the slots' scratch                               */
  int members;
} p3x_server_stats;

typedef struct p3x_server p3x_server;

void p3x_server_defaults(p3x_server_options* opts);
int p3x_server_create(p3x_context* ctx, const p3x_runtime* runtime, const p3x_server_options* opts,
                      p3x_server** server);
/* Register a member (its column set; infrastructure.dt is replaced per request);
 * allowed while serving.  P3X_ERR_BUDGET if its arrays do not fit the HBM budget
 * beside the members already registered, or its scratch can never fit one
 * slot's share of the scratch budget. */
int p3x_server_add_member(p3x_server* server, const p3x_column_set* set, int* member);
/* The neighbour step of a served process member (the mac_mic order of the
 * reference's multi-process runs): before P3, SHOC's post-process (when qc2 is
 * set: inv_qc_relvar of the process's fields from their qc and qc2, cldfrac_liq
 * clipped) and CldFraction (when cldfrac_liq is set: the fields' cldfrac_tot from
 * their qi and cldfrac_liq, plus the three other CldFraction outputs); around
 * it, the mass-and-energy column conservation check (when mass_err is set: the
 * column budgets before and after, in budget[4*ncol], and the relative errors
 * against the step's boundary fluxes, which the fields must then carry).  Every
 * array is device memory the caller owns, in the process's layout. */
typedef struct p3x_step_chain {
  const double* qc2;                        /* SHOC's qc variance; NULL = no SHOC stage        */
  double* cldfrac_liq;                      /* NULL = no CldFraction stage                     */
  double ice_threshold, ice_4out_threshold; /* CldFraction thresholds (1e-12, 1e-5 defaults)   */
  double *cldfrac_ice, *cldfrac_ice_4out, *cldfrac_tot_4out;
  double *mass_err, *energy_err;            /* ncol each; NULL = no conservation check        */
  double* budget;                           /* 4*ncol: mass and energy before, then after     */
  const double *u, *v, *ps, *phis;          /* optional inputs of the energy (NULL = 0)        */
} p3x_step_chain;
/* Register a member whose requests each run a whole P3Microphysics step (the
 * process's pre-amble, p3_main and post-amble, as p3x_process_run) on the
 * resident field-manager fields, with the optional neighbour stages of chain
 * (NULL = none).  The process object (p3x_process_create*, on the server's
 * context) holds the member's process-local arrays and must outlive the server;
 * one process per member.  Each request's results equal those of running the
 * same stages alone (the CldFraction, SHOC and budget entry points and
 * p3x_process_run) on one stream. */
int p3x_server_add_process(p3x_server* server, p3x_process* process, const p3x_atm_fields* fields,
                           const p3x_step_chain* chain, int* member);
/* Queue one step of member with time step dt; thread-safe. */
int p3x_server_submit(p3x_server* server, int member, double dt, unsigned long long tag);
/* Move up to max completed requests into out; waits up to wait_ms for one if
 * none is ready.  Returns the count (>= 0) or a negative P3X error. */
int p3x_server_poll(p3x_server* server, p3x_completion* out, int max, int wait_ms);
/* Wait until every submitted request has completed. */
int p3x_server_drain(p3x_server* server);
int p3x_server_stats_get(p3x_server* server, p3x_server_stats* stats);
void p3x_server_destroy(p3x_server* server);

/* ---- the ice lookup-table generator: the computation of
 * components/eam/tools/create_p3_lookupTable/create_p3_lookupTable_1.f90-v4.1
 * (the program that writes p3_lookup_table_1.dat, :170-1370) on the current HIP
 * device, and its write statements (:1374-1427).  Cells are ordered (rime
 * density 5, rime fraction 4, normalized ice mass 50); every value follows the
 * program's formula and summation order. ---- */
/* ice_rows[1000][15]: qsave nsave uns ums nagg nrwat vdep eff nlarge nsmall lsave
 * refl vdep1 dmm rhomm; coll_rows[1000][30][6]: qsave nsave lamrs nrrain qrrain
 * qsrain (host arrays).  elapsed_ms (optional): wall time of the generation. */
int p3x_ice_table_generate(double* ice_rows, double* coll_rows, void* stream, double* elapsed_ms);
/* The rows as a table file p3x_tables_load reads ("VERSION 4.1.1" header, the
 * program's e15.5 rows). */
int p3x_ice_table_write(const char* path, const double* ice_rows, const double* coll_rows);

const char* p3x_strerror(int status);

#ifdef __cplusplus
}
#endif
#endif /* P3X_H */
