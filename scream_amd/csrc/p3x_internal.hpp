// p3x_internal.hpp -- pieces of the C ABI implementation shared by its
// translation units (p3x_api.hip, p3x_server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/p3x.h"
#include "p3_launch.hpp"

// A context: device tables, grown-on-demand scratch, error word, timing events.
struct p3x_context {
  int device = 0;
  int ncu = 0;
  double* d_tab = nullptr;   // the device table set in FP64 (layout: TableLayout), one allocation
  float* d_tab_f = nullptr;  // the same in FP32 for the mixed-precision form
  double* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  unsigned* d_err = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_mid = nullptr;  // between the part 1 + 2 and the sedimentation kernels
  // the liquid sedimentation kernel's stream (beside the ice kernel) and its
  // fork and join events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipStream_t last_stream = nullptr;
  // the stream of the last p3x_main on this context (the last user of d_scratch);
  // ev1 marks the end of its work once main_launched is set
  hipStream_t main_stream = nullptr;
  int main_launched = 0;
  int pending_timing = 0;
  // device address of p3x_main_noalloc's fixed scratch array on this context's
  // device (NULL if it could not be resolved at p3x_init)
  double* noalloc_scratch = nullptr;
};

namespace p3x_detail {

int dev_fail(hipError_t e);
// The P3X status of a nonzero device error word.
int status_of_err(unsigned err);
// Default launch shape per (precision, nlev).
void default_shape(int precision, int nk, p3d::Shape* sh);
// Grows the context's scratch to at least need bytes (P3X_OK or an error status).
int grow_scratch(p3x_context* c, size_t need);
// Layout strides of a column set in k (P3X_ERR_ARG for an invalid layout or stride).
int set_geometry(int nj, int nk, int layout, int ld, int ld_iface, p3d::KParams* k);
// Elements a level field of k spans from its first element to its last.
size_t level_span(const p3d::KParams& k);
// p3x_main with (amble != NULL) the P3Microphysics pre/post-amble fused into its
// kernels; fixed_scratch != NULL: the step's scratch is that array of
// fixed_bytes bytes (the context's scratch is not touched or grown).
int main_impl(p3x_context* c, const p3x_runtime* runtime, const p3x_prognostic_state* ps,
              const p3x_diagnostic_inputs* di, const p3x_diagnostic_outputs* dout,
              const p3x_infrastructure* infra, const p3x_history_only* hist, int nj, int nk,
              const p3x_launch* opts, double* elapsed_us, const p3d::Amble* amble,
              double* fixed_scratch = nullptr, size_t fixed_bytes = 0);
// KParams of one column set; scratch (NULL: set later) points at SC_COUNT*nj*nk
// doubles followed by nj ints.  P3X_ERR_ARG for an invalid layout.
int fill_params(const p3x_context* c, const p3x_runtime* runtime, const p3x_prognostic_state* ps,
                const p3x_diagnostic_inputs* di, const p3x_diagnostic_outputs* dout,
                const p3x_infrastructure* infra, const p3x_history_only* hist, int nj, int nk, int layout,
                int ld, int ld_iface, int* col_status, int* sed_substeps, int* ice_table_idx, double* scratch,
                p3d::KParams* out);
// Every field pointer of a column set is non-null.
bool fields_present(const p3x_prognostic_state* ps, const p3x_diagnostic_inputs* di,
                    const p3x_diagnostic_outputs* dout, const p3x_history_only* hist);
// Bytes of scratch one column set of nj columns x nk levels needs, rounded up to 256.
size_t scratch_bytes_for(int nj, int nk);

// The p3_main arguments of one P3Microphysics process step (p3x_process.hip).
struct ProcessArgs {
  p3x_launch o;
  p3x_prognostic_state ps;
  p3x_diagnostic_inputs di;
  p3x_diagnostic_outputs dout;
  p3x_history_only hist;
  p3x_infrastructure inf;
  p3d::Amble amble;
};
int process_args(p3x_process* p, const p3x_atm_fields* f, double dt, const p3x_launch* opts, ProcessArgs* out);
// KParams of one process step (no scratch, no debug arrays): p3_main on the
// process's views with the pre/post-amble fused.
int process_kparams(p3x_process* p, const p3x_atm_fields* f, double dt, p3d::KParams* k);
// Bytes of the process's own device arrays, and its context.
size_t process_local_bytes(const p3x_process* p);
p3x_context* process_context(const p3x_process* p);
// The neighbour stages of a served process step (p3x_step_chain): argument
// check, and the stages before and after the step's kernels on stream s.
int chain_check(const p3x_process* p, const p3x_atm_fields* f, const p3x_step_chain* c);
int chain_before(const p3x_process* p, const p3x_atm_fields* f, const p3x_step_chain* c, hipStream_t s);
int chain_after(const p3x_process* p, const p3x_atm_fields* f, const p3x_step_chain* c, double dt, hipStream_t s);

// Launch shape for a batch whose deepest set has nk_max levels: the defaults,
// overridden by the nonzero options.  P3X_ERR_ARG if not compiled in.
int resolve_shape(int precision, int nk_max, int cb, int nt, int lpc, int wpe, p3d::Shape* sh);

// Host-side layout of a batch of column sets for one shape.
struct Layout {
  std::vector<int> first;      // part 1 virtual-block offsets, nsets + 1
  std::vector<int> first2;     // part 2 tile offsets, nsets + 1
  std::vector<int> sed_order;  // set indices grouped by sedimentation class (ascending L)
  std::vector<int> sed_first;  // per class nsets_c + 1 offsets, classes concatenated
  struct Cls {
    int L, set_off, first_off, nsets, total;
  };
  std::vector<Cls> cls;
};
int layout_batch(const p3d::KParams* sets, int n, const p3d::Shape& sh, Layout* out);
// Persistent grid for `total` virtual blocks: blocks_per_cu x CUs (0: the
// resident count from the occupancy API; -1: one block per virtual block).
int grid_for(int ncu, int blocks_per_cu, int total, int resident);
hipError_t occupancy(int precision, const p3d::Shape& sh, int L, int* r1, int* r2, int* rsed);
hipError_t launch_plan(int precision, const p3d::BatchPlan& plan, hipStream_t s);

}  // namespace p3x_detail
