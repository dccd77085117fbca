// C-ABI of the chemistry integrators (include/pele_chem_hip.h): handle
// management, reference defaults, host<->device staging for the 1-D form,
// kernel dispatch by mechanism, and the step-count reduction that gives the
// reference's int(average steps) return value.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pele_chem_hip.h"
#include "erk_tables.h"
#include "launch.h"

namespace ppc {
#define PPC_MECH(M) const MechLauncher* launcher_##M();
#include "inst/registry.inc"
#undef PPC_MECH
const MechLauncher* manifold_launcher(int dim);  // manifold.hip
}  // namespace ppc

namespace {

using ppc::MechLauncher;

const MechLauncher* find_mech(const char* name) {
  const MechLauncher* all[] = {
#define PPC_MECH(M) ppc::launcher_##M(),
#include "inst/registry.inc"
#undef PPC_MECH
  };
  for (const MechLauncher* m : all)
    if (name && std::strcmp(m->name, name) == 0) return m;
  return nullptr;
}

// "Manifold": the Manifold EOS over a table of dim dimensions (the reference's
// Eos_Model=Manifold with Chemistry_Model=Null and Manifold_Dim=dim).
bool is_manifold_name(const char* name) { return name && std::strcmp(name, "Manifold") == 0; }

const MechLauncher* resolve_mech(const char* name, int manifold_dim) {
  if (is_manifold_name(name)) return ppc::manifold_launcher(manifold_dim);
  return find_mech(name);
}

// Sums the per-cell counters and reduces min/max of the step counts.
__global__ void __launch_bounds__(256) reduce_stats(const int* __restrict__ nsteps, const int* __restrict__ nrhs,
                                                    const int* __restrict__ njac, const int* __restrict__ nlu,
                                                    const int* __restrict__ nnewton,
                                                    const int* __restrict__ nreject, int ncells,
                                                    unsigned long long* __restrict__ out) {
  // out: [0] sum steps, [1] rhs, [2] jac, [3] lu, [4] newton, [5] reject,
  //      [6] failed cells, [7] max steps, [8] min steps
  unsigned long long s[7] = {0, 0, 0, 0, 0, 0, 0};
  int mx = 0, mn = 0x7fffffff;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    const int v = nsteps[c];
    if (v < 0) {
      s[6] += 1;
    } else {
      s[0] += (unsigned long long)v;
      mx = v > mx ? v : mx;
      mn = v < mn ? v : mn;
    }
    if (nrhs) {
      s[1] += (unsigned long long)nrhs[c];
      s[2] += (unsigned long long)njac[c];
      s[3] += (unsigned long long)nlu[c];
      s[4] += (unsigned long long)nnewton[c];
      s[5] += (unsigned long long)nreject[c];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int i = 0; i < 7; ++i) s[i] += __shfl_down(s[i], off, 64);
    const int omx = __shfl_down(mx, off, 64);
    const int omn = __shfl_down(mn, off, 64);
    mx = omx > mx ? omx : mx;
    mn = omn < mn ? omn : mn;
  }
  if ((threadIdx.x & 63) == 0) {
    for (int i = 0; i < 7; ++i)
      if (s[i]) atomicAdd(&out[i], s[i]);
    atomicMax(&out[7], (unsigned long long)mx);
    atomicMin(&out[8], (unsigned long long)mn);
  }
}

// Counts the cells of the last launch that failed (negative step count) into
// *nfail, which the caller clears once per chain of calls.
__global__ void __launch_bounds__(256) count_failed(const int* __restrict__ nsteps, int ncells, int* __restrict__ nfail) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x)
    if (nsteps[c] < 0) atomicAdd(nfail, 1);
}

__global__ void init_stats(unsigned long long* out) {
  if (threadIdx.x < 9) out[threadIdx.x] = (threadIdx.x == 8) ? 0x7fffffffULL : 0ULL;
}

// ---- work-queue order: counting sort of the cells by the previous call's RHS
// count, most expensive first (longest-processing-time-first scheduling).
constexpr int NBINS = 256;

__global__ void zero_ints(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

PPC_DI int cost_bin(int nrhs) {
  const int b = nrhs / 2;
  return b < 0 ? 0 : (b > NBINS - 1 ? NBINS - 1 : b);
}

__global__ void __launch_bounds__(256) cost_hist(const int* __restrict__ nrhs, int ncells, int* __restrict__ bins) {
  __shared__ int h[NBINS];
  for (int i = threadIdx.x; i < NBINS; i += blockDim.x) h[i] = 0;
  __syncthreads();
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x)
    atomicAdd(&h[cost_bin(nrhs[c])], 1);
  __syncthreads();
  for (int i = threadIdx.x; i < NBINS; i += blockDim.x)
    if (h[i]) atomicAdd(&bins[i], h[i]);
}

// offs[b] = number of cells in bins above b (descending cost order); one block
__global__ void __launch_bounds__(NBINS) cost_scan(const int* __restrict__ bins, int* __restrict__ offs) {
  __shared__ int v[NBINS];
  const int b = threadIdx.x;
  v[b] = bins[NBINS - 1 - b];
  __syncthreads();
  for (int off = 1; off < NBINS; off <<= 1) {
    const int add = b >= off ? v[b - off] : 0;
    __syncthreads();
    v[b] += add;
    __syncthreads();
  }
  // v is the inclusive scan from the top bin down; exclusive offset of bin NBINS-1-b
  offs[NBINS - 1 - b] = v[b] - bins[NBINS - 1 - b];
}

__global__ void __launch_bounds__(256) cost_scatter(const int* __restrict__ nrhs, int ncells, int* __restrict__ offs,
                                                    int* __restrict__ order) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    const int pos = atomicAdd(&offs[cost_bin(nrhs[c])], 1);
    order[pos] = c;
  }
}

}  // namespace

// What a captured chained-call graph depends on.
struct GraphKey {
  const void* arrays[4];
  int ncells;
  double dt;
  const void *cnt, *order, *ws;
  long long ws_lanes;
  bool operator==(const GraphKey& o) const {
    return std::memcmp(arrays, o.arrays, sizeof(arrays)) == 0 && ncells == o.ncells && dt == o.dt &&
           cnt == o.cnt && order == o.order && ws == o.ws && ws_lanes == o.ws_lanes;
  }
};

struct pp_chem {
  pp_chem_config cfg;
  std::string mech_name;
  const MechLauncher* L = nullptr;
  ppc::ReactParamsDev P{};
  int device = 0;
  // per-cell counters (device), grown on demand
  int cap = 0;
  int* d_cnt = nullptr;  // 6 * cap ints
  unsigned long long* d_red = nullptr;
  // work queue of the batched forms: counter, visiting order, cost bins and
  // the identity of the last call (its per-cell costs order the next call)
  int* d_queue = nullptr;  // [0] counter, [1..NBINS] bins, [NBINS+1..2NBINS] offsets
  int* d_order = nullptr;  // cap ints
  bool hist_valid = false;
  int hist_ncells = 0;
  const void* hist_key = nullptr;
  // staging for host pointers in the 1-D form
  size_t stage_cap = 0;
  double* d_stage = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  pp_chem_stats stats{};
  std::string err;
  // adjoint: substep checkpoints (device), grown on demand
  size_t ckpt_cap = 0;
  double* d_ckpt = nullptr;
  // per-lane matrix workspace of the larger mechanisms, grown on demand
  size_t ws_cap = 0;
  double* d_ws = nullptr;
  // chained 1-D calls: one call's device work captured as a graph on a
  // stream of the handle, replayed; rebuilt when any captured pointer changes
  hipStream_t gstream = nullptr;
  hipEvent_t gev0 = nullptr, gev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  GraphKey gkey{};
  // persistent-grid bound of the batched forms: batched_waves_per_cu x CUs
  int max_blocks = 0;
  // failed cells over every call of a chain (pp_chem_react_1d_chained)
  int* d_fail = nullptr;
  // this device's address of the file-scope scratch array of
  // pp_chem_react_1d_noalloc (null when it could not be resolved)
  unsigned char* scratch = nullptr;
  // ReactorCvode: absolute tolerances atol x typical value per component
  double* d_atolv = nullptr;
  // Manifold EOS: the table's grids and values (device), set by
  // pp_chem_set_manifold
  bool manifold = false;
  double* d_man = nullptr;
  // ReactorArkode: the table (c, A, b, d) on the device
  double* d_erk = nullptr;
};

namespace {

int fail(pp_chem* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

// The Manifold EOS has nothing to evaluate before its table is set
// (Manifold.H:333-338 refuses to construct without it).
int manifold_unset(pp_chem* h) {
  return h->manifold && !h->P.man.values && !h->P.man.nn
             ? fail(h, PP_CHEM_EINVAL, "Manifold EOS must be initialized with table data (pp_chem_set_manifold)")
             : 0;
}

// What the Manifold EOS does not provide (Manifold.H: no Jacobian, no
// species thermodynamics; box_flatten constructs the EOS without its table).
int manifold_refuses(pp_chem* h, const char* what) {
  return h->manifold ? fail(h, PP_CHEM_EINVAL, std::string(what) + " is not available with the Manifold EOS") : 0;
}

int hip_fail(pp_chem* h, hipError_t e, const char* where) {
  return fail(h, PP_CHEM_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(h, x)                                   \
  do {                                                 \
    hipError_t e_ = (x);                               \
    if (e_ != hipSuccess) return hip_fail(h, e_, #x); \
  } while (0)

int ensure_counters(pp_chem* h, int ncells) {
  if (ncells <= h->cap) return 0;
  if (h->d_cnt) HIPCHK(h, hipFree(h->d_cnt));
  if (h->d_order) HIPCHK(h, hipFree(h->d_order));
  h->d_cnt = nullptr;
  h->d_order = nullptr;
  h->hist_valid = false;
  HIPCHK(h, hipMalloc(&h->d_cnt, sizeof(int) * 6 * (size_t)ncells));
  HIPCHK(h, hipMalloc(&h->d_order, sizeof(int) * (size_t)ncells));
  h->cap = ncells;
  return 0;
}

ppc::CellStats counters(pp_chem* h) {
  ppc::CellStats st;
  const size_t c = (size_t)h->cap;
  st.nsteps = h->d_cnt;
  st.nrhs = h->d_cnt + c;
  st.njac = h->d_cnt + 2 * c;
  st.nlu = h->d_cnt + 3 * c;
  st.nnewton = h->d_cnt + 4 * c;
  st.nreject = h->d_cnt + 5 * c;
  return st;
}

// Reduces the counters st of the last launch (into the 9-word buffer red) and
// returns int(average steps) or an error.
int finish_with(pp_chem* h, int ncells, hipStream_t s, float kernel_ms, const ppc::CellStats& st,
                unsigned long long* red);

int finish(pp_chem* h, int ncells, hipStream_t s, float kernel_ms) {
  return finish_with(h, ncells, s, kernel_ms, counters(h), h->d_red);
}

int finish_with(pp_chem* h, int ncells, hipStream_t s, float kernel_ms, const ppc::CellStats& st,
                unsigned long long* red) {
  hipLaunchKernelGGL(init_stats, dim3(1), dim3(64), 0, s, red);
  const int blocks = ncells > 0 ? (ncells + 255) / 256 < 1024 ? (ncells + 255) / 256 : 1024 : 1;
  hipLaunchKernelGGL(reduce_stats, dim3(blocks), dim3(256), 0, s, st.nsteps, st.nrhs, st.njac, st.nlu,
                     st.nnewton, st.nreject, ncells, red);
  HIPCHK(h, hipGetLastError());
  unsigned long long r[9];
  HIPCHK(h, hipMemcpyAsync(r, red, sizeof(r), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  pp_chem_stats& S = h->stats;
  S.ncells = ncells;
  S.nsteps = (long long)r[0];
  S.nrhs = (long long)r[1];
  S.njac = (long long)r[2];
  S.nlu = (long long)r[3];
  S.nnewton = (long long)r[4];
  S.nreject = (long long)r[5];
  S.nfailed = (int)r[6];
  S.max_steps_cell = (int)r[7];
  S.min_steps_cell = (int)(r[8] == 0x7fffffffULL ? 0 : r[8]);
  S.kernel_ms = kernel_ms;
  if (S.nfailed > 0)
    return fail(h, PP_CHEM_EINTEGRATE,
                std::to_string(S.nfailed) + " cell(s) failed to integrate (NaN, step limit or step too small)");
  if (ncells == 0) return 0;
  // ReactorCvode reports RHS evaluations (CVodeGetNumRhsEvals,
  // ReactorCvode.cpp:1365-1372), the other integrators steps
  // ReactorCvode and ReactorArkode report RHS evaluations (ReactorCvode.cpp:1365-1372,
  // ReactorArkode.cpp:239-260)
  const bool rhs_count = h->cfg.mode == PP_CHEM_CVODE || h->cfg.mode == PP_CHEM_ARKODE;
  return (int)((double)(rhs_count ? S.nrhs : S.nsteps) / (double)ncells);
}

bool is_batched(int mode) { return mode == ppc::MODE_BATCHED_MIXED || mode == ppc::MODE_BATCHED_FP64; }
// Kernels of a mode: ReactorCvode runs the plain FP64 adaptive BDF with its options.
int kmode(int mode) {
  return mode == PP_CHEM_CVODE ? ppc::MODE_ADAPTIVE : (mode == PP_CHEM_ARKODE ? ppc::MODE_ERK : mode);
}

// ---- device scratch of pp_chem_react_1d_noalloc: one array per GPU, held
// until the process ends.  Layout: [0, 256) queue counter, [256, 512) the
// statistics reduction, then the per-cell step counts, then the integrator
// workspace.
constexpr size_t kScratchBytes = 64ull << 20;
constexpr size_t kScratchCounts = 512;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Resolves this device's address of `scratch` once per device (the caller
// has made h->device current); host code uses that address, never the
// array's name.
hipError_t scratch_address(int device, unsigned char** out) {
  static std::mutex m;
  static unsigned char* addr[64] = {};
  std::lock_guard<std::mutex> lk(m);
  *out = nullptr;
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = static_cast<unsigned char*>(p);
  }
  *out = addr[device];
  return hipSuccess;
}

// Per-lane workspace of mechanisms too large for private memory: one row of
// doubles per lane and matrix element, for at most the resident waves per CU
// of the kernel build (4 SIMDs x 2 waves for the 2-wave batched build, x 1
// for the 1-wave batched build and the one-lane-per-cell kernels of the other
// integrators) and at most cfg.workspace_bytes (default kWsMaxBytes).
constexpr long long kWsMaxBytes = 32LL << 30;

int prepare_workspace(pp_chem* h, int ncells) {
  h->P.ws = nullptr;
  h->P.ws_lanes = 0;
  const bool group = is_batched(h->cfg.mode) && h->cfg.batched_cells_per_wave != 0;
  const long long rows = group ? h->L->group_ws_rows(h->cfg.batched_cells_per_wave) : h->L->ws_rows(kmode(h->cfg.mode));
  if (rows == 0 || ncells == 0) return 0;
  int ncu = 0;
  HIPCHK(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
  const long long per_cu =
      (is_batched(h->cfg.mode) && (group || h->cfg.batched_waves_per_simd != 1)) ? 8 : 4;
  // blocks of 64 lanes that can hold a cell each: one cell per lane, or per
  // group of 64 / batched_cells_per_wave lanes
  const long long lanes_per_cell = group ? 64 / h->cfg.batched_cells_per_wave : 1;
  long long blocks = std::min<long long>(((long long)ncells * lanes_per_cell + 63) / 64, per_cu * std::max(ncu, 1));
  if (h->max_blocks > 0) blocks = std::min<long long>(blocks, h->max_blocks);
  const long long cap = h->cfg.workspace_bytes > 0 ? h->cfg.workspace_bytes : kWsMaxBytes;
  const long long lane_bytes = 64 * rows * (long long)sizeof(double);  // one 64-lane block
  if (lane_bytes > cap) return fail(h, PP_CHEM_EINVAL, "workspace_bytes below one 64-cell block");
  blocks = std::min(blocks, cap / lane_bytes);
  const size_t need = (size_t)(blocks * 64 * rows);
  if (need > h->ws_cap) {
    if (h->d_ws) HIPCHK(h, hipFree(h->d_ws));
    h->d_ws = nullptr;
    h->ws_cap = 0;
    HIPCHK(h, hipMalloc(&h->d_ws, need * sizeof(double)));
    h->ws_cap = need;
  }
  h->P.ws = h->d_ws;
  h->P.ws_lanes = blocks * 64;
  return 0;
}

// Prepares the work queue: zero the counter and, when the previous call
// integrated the same cells, order them by that call's RHS counts.
ppc::QueueArgs prepare_queue(pp_chem* h, int ncells, const void* key, hipStream_t s) {
  ppc::QueueArgs q{nullptr, h->d_queue, ncells, h->max_blocks};
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(256), 0, s, h->d_queue, 1 + 2 * NBINS);
  if (h->hist_valid && h->hist_ncells == ncells && h->hist_key == key && ncells > 0) {
    int* bins = h->d_queue + 1;
    int* offs = bins + NBINS;
    const int blocks = std::min(1024, (ncells + 255) / 256);
    hipLaunchKernelGGL(cost_hist, dim3(blocks), dim3(256), 0, s, h->d_cnt + (size_t)h->cap, ncells, bins);
    hipLaunchKernelGGL(cost_scan, dim3(1), dim3(NBINS), 0, s, bins, offs);
    hipLaunchKernelGGL(cost_scatter, dim3(blocks), dim3(256), 0, s, h->d_cnt + (size_t)h->cap, ncells, offs,
                       h->d_order);
    q.order = h->d_order;
  }
  h->hist_valid = true;
  h->hist_ncells = ncells;
  h->hist_key = key;
  return q;
}

// One 1-D call's device work on stream s (the queue in cost order when the
// previous call integrated the same cells).  With count_fail, the cells that
// failed in this call are added to h->d_fail.
hipError_t enqueue_1d(pp_chem* h, double* y, double* ysrc, double* rX, double* rXsrc, double dt, int ncells,
                      const void* key, hipStream_t s, bool count_fail = false) {
  hipError_t e;
  if (is_batched(h->cfg.mode)) {
    const ppc::QueueArgs q = prepare_queue(h, ncells, key, s);
    e = h->L->batched_1d(h->cfg.mode, h->P, y, ysrc, rX, rXsrc, dt, q, counters(h), s);
  } else {
    e = h->L->react_1d(kmode(h->cfg.mode), h->P, y, ysrc, rX, rXsrc, dt, ncells, counters(h), h->cfg.block_size, s);
  }
  if (e != hipSuccess || !count_fail || ncells == 0) return e;
  const int blocks = std::min(1024, (ncells + 255) / 256);
  hipLaunchKernelGGL(count_failed, dim3(blocks), dim3(256), 0, s, h->d_cnt, ncells, h->d_fail);
  return hipGetLastError();
}

void drop_graph(pp_chem* h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
}

ppc::ReactParamsDev to_dev(const pp_chem_config& c) {
  ppc::ReactParamsDev P{};
  P.reactor_type = c.reactor_type;
  P.bdf = {c.bdf_nsubsteps, c.bdf_nonlinear_iters, c.bdf_gmres_restarts, c.bdf_gmres_kspiters,
           c.bdf_gmres_precond, c.bdf_scheme, c.bdf_nonlinear_tol, c.bdf_gmres_tol};
  P.rk64 = {c.rk64_atol, c.rk64_nsubsteps_guess, c.rk64_nsubsteps_min, c.rk64_nsubsteps_max};
  P.adapt = {c.rtol, c.atol, c.first_step, c.max_steps, c.batched_lu_batch, c.batched_lu_wait,
             c.batched_waves_per_simd, c.batched_cells_per_wave, c.cvode_max_order, c.max_nls_iters,
             c.jac_max_age, (c.mode == PP_CHEM_CVODE || c.mode == PP_CHEM_ARKODE) ? 1 : 0, nullptr};
  return P;
}

ppc::A4 to_a4(const pp_array4* a) {
  ppc::A4 r{};
  if (!a || !a->p) return r;
  r.p = a->p;
  r.js = a->jstride;
  r.ks = a->kstride;
  r.ns = a->nstride;
  r.bx = a->begin[0];
  r.by = a->begin[1];
  r.bz = a->begin[2];
  return r;
}

// Box geometry and views; every non-null view must cover the box and have
// enough components.
int box_args(pp_chem* h, const int lo[3], const int hi[3], const pp_array4* rY, const pp_array4* rYsrc,
             const pp_array4* T, const pp_array4* rE, const pp_array4* rEsrc, const pp_array4* FC,
             ppc::BoxArgs& b, int& ncells) {
  long long nc = 1;
  b = ppc::BoxArgs{};
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = lo[d];
    b.len[d] = hi[d] - lo[d] + 1;
    if (b.len[d] < 0) return fail(h, PP_CHEM_EINVAL, "hi < lo");
    nc *= b.len[d];
  }
  if (nc > 0x7fffffffLL) return fail(h, PP_CHEM_EINVAL, "box too large");
  ncells = (int)nc;
  const pp_array4* views[] = {rY, rYsrc, T, rE, rEsrc, FC};
  const int need[] = {h->L->ns, h->L->ns, 1, 1, 1, 1};
  for (int v = 0; v < 6; ++v) {
    const pp_array4* a = views[v];
    if (!a || !a->p) continue;
    for (int d = 0; d < 3; ++d)
      if (lo[d] < a->begin[d] || hi[d] >= a->end[d]) return fail(h, PP_CHEM_EINVAL, "Array4 does not cover the box");
    if (a->ncomp < need[v]) return fail(h, PP_CHEM_EINVAL, "Array4 has too few components");
  }
  b.rY = to_a4(rY);
  b.rYsrc = to_a4(rYsrc);
  b.T = to_a4(T);
  b.rE = to_a4(rE);
  b.rEsrc = to_a4(rEsrc);
  b.FC = to_a4(FC);
  return 0;
}

}  // namespace

extern "C" {

int pp_chem_list_mechanisms(char* buf, int buflen) {
  std::string s;
#define PPC_MECH(M) s += (s.empty() ? "" : ",") + std::string(#M);
#include "inst/registry.inc"
#undef PPC_MECH
  s += ",Manifold";
  if (buf && buflen > 0) {
    std::strncpy(buf, s.c_str(), (size_t)buflen - 1);
    buf[buflen - 1] = '\0';
  }
  return (int)s.size() + 1;
}

int pp_chem_config_init(pp_chem_config* cfg, const char* mechanism, int mode, int reactor_type) {
  if (!cfg) return PP_CHEM_EINVAL;
  const MechLauncher* L = resolve_mech(mechanism, 1);
  if (!L) return PP_CHEM_EMECH;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->mechanism = L->name;
  // MANIFOLD_DIM defaults to 1 (Source/Eos/Manifold.H:13-15)
  if (is_manifold_name(mechanism)) cfg->manifold_dim = 1;
  cfg->reactor_type = reactor_type;
  cfg->mode = mode;
  cfg->mod_reactor = 0;
  // ReactorBDF defaults (ReactorBDF.H:118-128)
  cfg->bdf_nsubsteps = 1;
  cfg->bdf_nonlinear_iters = 1;
  cfg->bdf_gmres_restarts = 10;
  cfg->bdf_gmres_kspiters = L->ns + 1;
  cfg->bdf_gmres_precond = 2;
  cfg->bdf_scheme = 3;
  cfg->bdf_nonlinear_tol = 1e-6;
  cfg->bdf_gmres_tol = 1e-10;
  // ReactorRK64 defaults (ReactorRK64.H:131-134)
  cfg->rk64_atol = 1e-10;
  cfg->rk64_nsubsteps_guess = 10;
  cfg->rk64_nsubsteps_min = 5;
  cfg->rk64_nsubsteps_max = 500;
  // adaptive BDF
  cfg->rtol = 1e-8;
  cfg->atol = 1e-10;
  cfg->first_step = 0.0;
  cfg->max_steps = 100000;
  cfg->block_size = 0;
  if (mode == PP_CHEM_CVODE) {
    // ReactorCvode defaults (ReactorCvode.H:179-192, ReactorCvode.cpp:214-218):
    // rtol 1e-6, atol 1e-10, cvode.max_order 2, cvode.max_substeps 10000,
    // ode.max_nls_iters 3, Jacobian re-evaluated after 100 steps
    cfg->rtol = 1e-6;
    cfg->atol = 1e-10;
    cfg->max_steps = 10000;
    cfg->cvode_max_order = 2;
    cfg->max_nls_iters = 3;
    cfg->jac_max_age = 100;
  }
  if (mode == PP_CHEM_ARKODE) {
    // ReactorArkode defaults (ReactorArkode.H:132-136): rtol 1e-6, atol 1e-10,
    // ARKODE_ZONNEVELD_5_3_4 with the PID controller; ARKODE's 500 steps per call
    cfg->rtol = 1e-6;
    cfg->atol = 1e-10;
    cfg->max_steps = 500;
    cfg->erk_method = 40;
    cfg->erk_controller = 0;
  }
  return PP_CHEM_OK;
}

int pp_chem_create(const pp_chem_config* cfg, pp_chem** out) {
  if (!cfg || !out) return PP_CHEM_EINVAL;
  *out = nullptr;
  const bool manifold = is_manifold_name(cfg->mechanism);
  if (manifold && (cfg->manifold_dim < 1 || cfg->manifold_dim > 5)) return PP_CHEM_EINVAL;
  const MechLauncher* L = resolve_mech(cfg->mechanism, cfg->manifold_dim);
  if (!L) return PP_CHEM_EMECH;
  // the Manifold EOS is propagated through ReactorRK64 only (ReactorBase.H:84-97)
  if (manifold && cfg->mode != PP_CHEM_RK64) return PP_CHEM_EINVAL;
  // ReactorTypes::check_reactor_type (ReactorTypes.H:12-19) as an error code
  if (cfg->reactor_type != 1 && cfg->reactor_type != 2) return PP_CHEM_EINVAL;
  if (cfg->mode < PP_CHEM_BDF || cfg->mode > PP_CHEM_ARKODE) return PP_CHEM_EINVAL;
  // ReactorArkode's table and controller (ReactorArkode.cpp:26-110); the
  // ARK437L2SA and ARK548L2SAb tables are not provided
  if (cfg->mode == PP_CHEM_ARKODE) {
    if (cfg->erk_method == 43 || cfg->erk_method == 54) return PP_CHEM_EINVAL;
    if (cfg->erk_controller < 0 || cfg->erk_controller > 3) return PP_CHEM_EINVAL;
  } else if (cfg->erk_method != 0 || cfg->erk_controller != 0) {
    return PP_CHEM_EINVAL;
  }
  // CVODE-style options: the plain FP64 adaptive BDF and ReactorCvode only
  if (cfg->cvode_max_order < 0 || cfg->cvode_max_order > 5 || cfg->max_nls_iters < 0 || cfg->jac_max_age < 0)
    return PP_CHEM_EINVAL;
  if ((cfg->cvode_max_order || cfg->max_nls_iters || cfg->jac_max_age) && cfg->mode != PP_CHEM_BDF_ADAPTIVE &&
      cfg->mode != PP_CHEM_CVODE)
    return PP_CHEM_EINVAL;
  if (is_batched(cfg->mode) && !L->batched_ok) return PP_CHEM_EINVAL;  // form not compiled for L
  if (cfg->mode == PP_CHEM_BDF &&
      (cfg->bdf_nsubsteps < 1 || cfg->bdf_scheme < 0 || cfg->bdf_scheme > 3 ||
       cfg->bdf_gmres_precond < 0 || cfg->bdf_gmres_precond > 2 || cfg->bdf_gmres_kspiters < 1 ||
       cfg->bdf_gmres_kspiters > L->ns + 1 || cfg->bdf_nonlinear_iters < 1))
    return PP_CHEM_EINVAL;
  if (cfg->mode == PP_CHEM_RK64 &&
      (cfg->rk64_nsubsteps_guess < 1 || cfg->rk64_nsubsteps_min < 1 || cfg->rk64_nsubsteps_max < 1))
    return PP_CHEM_EINVAL;
  if (cfg->mode >= PP_CHEM_BDF_ADAPTIVE &&
      (!(cfg->rtol > 0) || !(cfg->atol > 0) || cfg->max_steps < 1))
    return PP_CHEM_EINVAL;
  if (cfg->workspace_bytes < 0) return PP_CHEM_EINVAL;
  if (cfg->batched_waves_per_cu < 0 || cfg->batched_lu_batch < 0 || cfg->batched_lu_batch > 64 ||
      cfg->batched_lu_wait < 0)
    return PP_CHEM_EINVAL;
  if (cfg->clean_init_massfrac != 0 && cfg->clean_init_massfrac != 1) return PP_CHEM_EINVAL;
  if (cfg->batched_box_direct != 0 && cfg->batched_box_direct != 1) return PP_CHEM_EINVAL;
  // the kernel build: batched forms 0, 1 or 2 waves per SIMD; the plain
  // adaptive BDF 0 (the compiler's choice), or 2 or 4 for mechanisms whose
  // cells fit in registers (no workspace); the parity integrators 0
  const int occ = cfg->batched_waves_per_simd;
  if (is_batched(cfg->mode) ? (occ < 0 || occ > 2)
                            : ((cfg->mode == PP_CHEM_BDF_ADAPTIVE || cfg->mode == PP_CHEM_CVODE)
                                   ? (occ != 0 && ((occ != 2 && occ != 4) || L->ws_rows(kmode(cfg->mode)) > 0))
                                   : occ != 0))
    return PP_CHEM_EINVAL;
  if (cfg->batched_waves_per_simd == 1 && cfg->batched_box_direct == 1) return PP_CHEM_EINVAL;  // no such build
  if (cfg->batched_cells_per_wave != 0) {
    // a group kernel of this size must exist for the mechanism; box calls go through flat state
    if (!is_batched(cfg->mode) || cfg->batched_box_direct == 1) return PP_CHEM_EINVAL;
    if (L->group_ws_rows(cfg->batched_cells_per_wave) <= 0) return PP_CHEM_EINVAL;
  }
  const int bs = cfg->block_size > 0 ? cfg->block_size : 64;
  if (bs > 64 || bs % 64 != 0) return PP_CHEM_EINVAL;  // kernels are built for 64-lane blocks
  pp_chem* h = new pp_chem();
  h->cfg = *cfg;
  h->mech_name = L->name;
  h->cfg.mechanism = h->mech_name.c_str();
  h->cfg.block_size = bs;
  h->L = L;
  h->manifold = manifold;
  h->P = to_dev(*cfg);
  if (cfg->mode == PP_CHEM_ARKODE) {
    // unknown codes select ARKODE_ZONNEVELD_5_3_4, as ReactorArkode.cpp:78-81 does
    const ppc::ErkTableHost* T = nullptr;
    for (const auto& t : ppc::ERK_TABLES)
      if (t.code == cfg->erk_method) T = &t;
    for (const auto& t : ppc::ERK_TABLES)
      if (!T && t.code == ppc::ERK_DEFAULT) T = &t;
    const int s = T->s;
    std::vector<double> tab;
    tab.insert(tab.end(), T->c, T->c + s);
    tab.insert(tab.end(), T->A, T->A + s * s);
    tab.insert(tab.end(), T->b, T->b + s);
    tab.insert(tab.end(), T->d, T->d + s);
    hipError_t et = hipMalloc(&h->d_erk, sizeof(double) * tab.size());
    if (et == hipSuccess) et = hipMemcpy(h->d_erk, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
    if (et != hipSuccess) {
      pp_chem_destroy(h);
      return PP_CHEM_EHIP;
    }
    h->P.erk = {h->d_erk, s, T->q, T->p, T->fsal, cfg->erk_controller};
    h->cfg.erk_method = T->code;
  }
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess && cfg->batched_waves_per_cu > 0) {
    int ncu = 0;
    e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device);
    h->max_blocks = cfg->batched_waves_per_cu * ncu;  // one 64-lane wave per block
  }
  if (e == hipSuccess) e = hipMalloc(&h->d_red, sizeof(unsigned long long) * 9);
  if (e == hipSuccess) e = hipMalloc(&h->d_queue, sizeof(int) * (1 + 2 * NBINS));
  if (e == hipSuccess) e = hipMalloc(&h->d_fail, sizeof(int));
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e != hipSuccess) {
    pp_chem_destroy(h);
    return PP_CHEM_EHIP;
  }
  // the device scratch of pp_chem_react_1d_noalloc: a failure here only
  // disables that entry point (it reports the error)
  if (scratch_address(h->device, &h->scratch) != hipSuccess) h->scratch = nullptr;
  *out = h;
  return PP_CHEM_OK;
}

void pp_chem_destroy(pp_chem* h) {
  if (!h) return;
  drop_graph(h);
  if (h->gstream) (void)hipStreamDestroy(h->gstream);
  if (h->gev0) (void)hipEventDestroy(h->gev0);
  if (h->gev1) (void)hipEventDestroy(h->gev1);
  if (h->d_ckpt) (void)hipFree(h->d_ckpt);
  if (h->d_ws) (void)hipFree(h->d_ws);
  if (h->d_cnt) (void)hipFree(h->d_cnt);
  if (h->d_red) (void)hipFree(h->d_red);
  if (h->d_queue) (void)hipFree(h->d_queue);
  if (h->d_fail) (void)hipFree(h->d_fail);
  if (h->d_order) (void)hipFree(h->d_order);
  if (h->d_stage) (void)hipFree(h->d_stage);
  if (h->d_atolv) (void)hipFree(h->d_atolv);
  if (h->d_man) (void)hipFree(h->d_man);
  if (h->d_erk) (void)hipFree(h->d_erk);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  delete h;
}

int pp_chem_set_typ_vals(pp_chem* h, const double* typ_vals, int n) {
  if (!h) return PP_CHEM_EINVAL;
  if (!typ_vals || n != h->L->ns + 1) return fail(h, PP_CHEM_EINVAL, "typ_vals needs NUM_SPECIES + 1 values");
  // read by ReactorCvode only, as the reference's ReactorBDF / ReactorRK64
  // ignore them; typ_vals[0] <= 0 restores the scalar atol
  if (h->cfg.mode != PP_CHEM_CVODE && h->cfg.mode != PP_CHEM_ARKODE) return PP_CHEM_OK;
  if (!(typ_vals[0] > 0.0)) {
    h->P.adapt.atolv = nullptr;
    drop_graph(h);
    return PP_CHEM_OK;
  }
  std::vector<double> a((size_t)n);
  for (int k = 0; k < n; ++k) {
    if (!(typ_vals[k] > 0.0) || !std::isfinite(typ_vals[k]))
      return fail(h, PP_CHEM_EINVAL, "typical values must be positive and finite");
    a[(size_t)k] = typ_vals[k] * h->cfg.atol;
  }
  HIPCHK(h, hipSetDevice(h->device));
  if (!h->d_atolv) HIPCHK(h, hipMalloc(&h->d_atolv, sizeof(double) * (size_t)n));
  HIPCHK(h, hipMemcpy(h->d_atolv, a.data(), sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
  h->P.adapt.atolv = h->d_atolv;
  drop_graph(h);
  return PP_CHEM_OK;
}

int pp_chem_set_manifold(pp_chem* h, const pp_chem_manifold* m) {
  if (!h) return PP_CHEM_EINVAL;
  if (!h->manifold) return fail(h, PP_CHEM_EINVAL, "manifold parameters need the Manifold EOS (mechanism \"Manifold\")");
  if (!m) return fail(h, PP_CHEM_EINVAL, "null argument");
  if (m->model != 0 && m->model != 1) return fail(h, PP_CHEM_EINVAL, "manifold model must be 0 (table) or 1 (network)");
  const bool net = m->model == 1;
  if (net ? (!m->nn_kind || !m->nn_nout || !m->nn_params) : (!m->dim_lengths || !m->grids || !m->values))
    return fail(h, PP_CHEM_EINVAL, "null argument");
  if (m->ndim != h->cfg.manifold_dim)
    return fail(h, PP_CHEM_EINVAL,
                net ? "network inputs differ from manifold_dim (NUM_SPECIES - 1)"
                    : "table dimension differs from manifold_dim (NUM_SPECIES - 1)");
  if (m->nvar < 1) return fail(h, PP_CHEM_EINVAL, "no tabulated variables");
  ppc::ManifoldDev d{};
  d.ndim = m->ndim;
  d.nvar = m->nvar;
  d.model = m->model;
  long long ngrid = 0, stride = 1, nparam = 0;
  if (net) {
    if (m->nn_layers < 1 || m->nn_layers > ppc::MAN_NN_MAXL)
      return fail(h, PP_CHEM_EINVAL, "a network has 1 to 16 layers");
    int width = m->ndim;
    for (int l = 0; l < m->nn_layers; ++l) {
      const int kind = m->nn_kind[l];
      int nout = width;
      long long np;
      if (kind == 0) {
        nout = m->nn_nout[l];
        np = (long long)width * nout + nout;
      } else if (kind == 1) {
        np = 1;
      } else if (kind == 2) {
        np = 2LL * width;
      } else {
        return fail(h, PP_CHEM_EINVAL, "network layer kind must be 0 (Linear), 1 (LeakyReLU) or 2 (BatchNorm1d)");
      }
      if (nout < 1 || nout >= ppc::MAN_NN_MAXW)
        return fail(h, PP_CHEM_EINVAL, "network layers must have 1 to 99 outputs (MAXD_NETWORK)");
      d.nn_kind[l] = kind;
      d.nn_nin[l] = width;
      d.nn_nout[l] = nout;
      d.nn_off[l] = (int)nparam;
      nparam += np;
      width = nout;
    }
    if (width != m->nvar) return fail(h, PP_CHEM_EINVAL, "the network's last layer must give nvar outputs");
    d.nn_layers = m->nn_layers;
  } else {
    for (int k = 0; k < m->ndim; ++k) {
      const int n = m->dim_lengths[k];
      if (n < 2) return fail(h, PP_CHEM_EINVAL, "a table dimension has fewer than 2 points");
      for (int i = 0; i + 1 < n; ++i)
        if (!(m->grids[ngrid + i + 1] > m->grids[ngrid + i]))
          return fail(h, PP_CHEM_EINVAL, "table grids must be strictly ascending");
      d.dim_len[k] = n;
      d.dim_off[k] = (int)ngrid;
      d.dim_stride[k] = (int)stride;
      ngrid += n;
      stride *= n;
      if (stride > 0x7fffffffLL / 8) return fail(h, PP_CHEM_EINVAL, "table too large");
    }
    d.var_stride = (int)stride;
  }
  const auto var_ok = [&](int v, bool opt) { return (opt && v == -1) || (v >= 0 && v < m->nvar); };
  if (!var_ok(m->idx_density, false) || !var_ok(m->idx_T, false))
    return fail(h, PP_CHEM_EINVAL, "density or temperature variable index out of range");
  for (int k = 0; k < m->ndim; ++k)
    if (!var_ok(m->idx_src[k], true)) return fail(h, PP_CHEM_EINVAL, "source variable index out of range");
  if (m->density_lookup < 0 || m->density_lookup > 2) return fail(h, PP_CHEM_EINVAL, "bad density lookup type");
  d.idx_density = m->idx_density;
  d.dens_lookup = m->density_lookup;
  d.idx_T = m->idx_T;
  d.compute_T = m->compute_temperature ? 1 : 0;
  for (int k = 0; k < ppc::MAN_MAXD; ++k) d.idx_src[k] = k < m->ndim ? m->idx_src[k] : -1;
  d.pnom = m->nominal_pressure;
  const size_t nval = net ? (size_t)nparam : (size_t)m->nvar * (size_t)stride;
  HIPCHK(h, hipSetDevice(h->device));
  if (h->d_man) HIPCHK(h, hipFree(h->d_man));
  h->d_man = nullptr;
  h->P.man = ppc::ManifoldDev{};
  HIPCHK(h, hipMalloc(&h->d_man, sizeof(double) * ((size_t)ngrid + nval)));
  if (net) {
    HIPCHK(h, hipMemcpy(h->d_man, m->nn_params, sizeof(double) * nval, hipMemcpyHostToDevice));
    d.nn = h->d_man;
  } else {
    HIPCHK(h, hipMemcpy(h->d_man, m->grids, sizeof(double) * (size_t)ngrid, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_man + ngrid, m->values, sizeof(double) * nval, hipMemcpyHostToDevice));
    d.grids = h->d_man;
    d.values = h->d_man + ngrid;
  }
  h->P.man = d;
  drop_graph(h);
  return PP_CHEM_OK;
}

int pp_chem_num_species(const pp_chem* h) { return h ? h->L->ns : PP_CHEM_EINVAL; }
int pp_chem_num_reactions(const pp_chem* h) { return h ? h->L->nr : PP_CHEM_EINVAL; }
long long pp_chem_workspace_rows(const pp_chem* h) {
  if (!h) return PP_CHEM_EINVAL;
  if (is_batched(h->cfg.mode) && h->cfg.batched_cells_per_wave != 0)
    return h->L->group_ws_rows(h->cfg.batched_cells_per_wave);
  return h->L->ws_rows(kmode(h->cfg.mode));
}

int pp_chem_adjoint_1d(pp_chem* h, double* y, const double* rX, int ncells, int nmacro, double dt,
                       int spec_a, int spec_b, double* out, double* grad, void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_refuses(h, "the adjoint")) return r;
  if (ncells < 0 || nmacro < 1 || !(dt > 0)) return fail(h, PP_CHEM_EINVAL, "bad argument");
  if (ncells > 0 && (!y || !rX || !out || !grad)) return fail(h, PP_CHEM_EINVAL, "null array");
  if (spec_a < 0 || spec_a >= h->L->ns || spec_b < 0 || spec_b >= h->L->ns)
    return fail(h, PP_CHEM_EINVAL, "output species index out of range");
  if (h->cfg.bdf_nonlinear_iters != 1)
    return fail(h, PP_CHEM_EINVAL, "the adjoint differentiates one Newton iteration per substep");
  // the adjoint runs on the table-driven form of the mechanism
  const MechLauncher* L = h->L;
  if (L->nr > 128) return fail(h, PP_CHEM_EINVAL, "more than 128 reactions");
  if (L->ns + 1 > 64) return fail(h, PP_CHEM_EINVAL, "more than 63 species");
  hipStream_t s = (hipStream_t)stream;
  if (ncells == 0) return 0;
  const std::string tab = std::string(L->name) + "_table";
  const MechLauncher* LA = find_mech(tab.c_str());
  if (!LA) LA = L;
  const size_t need = (size_t)ncells * (size_t)nmacro * (size_t)h->cfg.bdf_nsubsteps * (size_t)(L->ns + 1);
  if (need > h->ckpt_cap) {
    if (h->d_ckpt) HIPCHK(h, hipFree(h->d_ckpt));
    h->d_ckpt = nullptr;
    HIPCHK(h, hipMalloc(&h->d_ckpt, need * sizeof(double)));
    h->ckpt_cap = need;
  }
  ppc::AdjointArgs a{y, rX, grad, out, h->d_ckpt, ncells, nmacro, spec_a, spec_b, dt};
  HIPCHK(h, hipEventRecord(h->ev0, s));
  const hipError_t e = LA->adjoint_1d(h->P, a, s);
  if (e == hipErrorNotSupported) return fail(h, PP_CHEM_EINVAL, "no table-driven form of " + std::string(L->name));
  if (e != hipSuccess) return hip_fail(h, e, "adjoint launch");
  HIPCHK(h, hipEventRecord(h->ev1, s));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->stats.kernel_ms = ms;
  h->stats.ncells = ncells;
  return 0;
}

int pp_chem_react_box(pp_chem* h, const int lo[3], const int hi[3], const pp_array4* rY,
                      const pp_array4* rYsrc, const pp_array4* T, const pp_array4* rE,
                      const pp_array4* rEsrc, const pp_array4* FC, const pp_array4* mask,
                      double* dt_react, double* time, void* stream) {
  (void)mask;  // ignored by the reference (ReactorBDF.cpp:322)
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_unset(h)) return r;
  if (!lo || !hi || !rY || !rYsrc || !T || !rE || !rEsrc || !dt_react || !time)
    return fail(h, PP_CHEM_EINVAL, "null argument");
  if (!rY->p || !rYsrc->p || !T->p || !rE->p || !rEsrc->p) return fail(h, PP_CHEM_EINVAL, "null array");
  ppc::BoxArgs b{};
  int ncells = 0;
  if (int r = box_args(h, lo, hi, rY, rYsrc, T, rE, rEsrc, FC, b, ncells)) return r;
  if (int r = ensure_counters(h, ncells > 0 ? ncells : 1)) return r;
  if (int r = prepare_workspace(h, ncells)) return r;
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(h, hipEventRecord(h->ev0, s));
  if (ncells > 0) {
    hipError_t e;
    if (is_batched(h->cfg.mode) && !h->cfg.batched_box_direct) {
      // flat state, the 1-D kernel, back to the box: the same operations per
      // cell as the box kernel (T from e / rho with the box T as guess;
      // rE + dt rEsrc; T from the new energy with the integrated T as guess;
      // FC = steps), and the 1-D kernel is the faster build of the two
      const int ns = h->L->ns;
      const size_t nc = (size_t)ncells, need = nc * (size_t)(2 * ns + 3);
      if (need > h->stage_cap) {
        if (h->d_stage) HIPCHK(h, hipFree(h->d_stage));
        h->d_stage = nullptr;
        h->stage_cap = 0;
        HIPCHK(h, hipMalloc(&h->d_stage, need * sizeof(double)));
        h->stage_cap = need;
      }
      ppc::FlatArgs fa{};
      fa.b = b;
      fa.y = h->d_stage;
      fa.ysrc = fa.y + nc * (ns + 1);
      fa.e = fa.ysrc + nc * ns;
      fa.esrc = fa.e + nc;
      fa.reactor_type = h->cfg.reactor_type;
      fa.divide = 1;
      e = h->L->flatten(fa, 0, s);
      if (e == hipSuccess) {
        const ppc::QueueArgs q = prepare_queue(h, ncells, rY->p, s);
        e = h->L->batched_1d(h->cfg.mode, h->P, fa.y, fa.ysrc, fa.e, fa.esrc, *dt_react, q, counters(h), s);
      }
      if (e == hipSuccess) {
        fa.dt = 0.0;  // the 1-D kernel has already advanced e by dt * rEsrc
        fa.fc32 = counters(h).nsteps;
        e = h->L->flatten(fa, 1, s);
      }
    } else if (is_batched(h->cfg.mode)) {
      const ppc::QueueArgs q = prepare_queue(h, ncells, rY->p, s);
      e = h->L->batched_box(h->cfg.mode, h->P, b, *dt_react, q, counters(h), s);
    } else {
      e = h->L->react_box(kmode(h->cfg.mode), h->P, b, *dt_react, ncells, counters(h), h->cfg.block_size, s);
    }
    if (e != hipSuccess) return hip_fail(h, e, "react_box launch");
  }
  HIPCHK(h, hipEventRecord(h->ev1, s));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  const int r = finish(h, ncells, s, ms);
  if (h->cfg.mod_reactor) *time += *dt_react;
  return r;
}

int pp_chem_react_1d(pp_chem* h, double* y, double* ysrc, double* rX, double* rXsrc, double* dt_react,
                     double* time, int ncells, int on_device, void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_unset(h)) return r;
  if (!dt_react || !time || ncells < 0) return fail(h, PP_CHEM_EINVAL, "bad argument");
  if (ncells > 0 && (!y || !ysrc || !rX || !rXsrc)) return fail(h, PP_CHEM_EINVAL, "null array");
  const int ns = h->L->ns, n = ns + 1;
  hipStream_t s = (hipStream_t)stream;
  if (int r = ensure_counters(h, ncells > 0 ? ncells : 1)) return r;
  if (int r = prepare_workspace(h, ncells)) return r;
  double *dy = y, *dys = ysrc, *dx = rX, *dxs = rXsrc;
  const size_t ny = (size_t)ncells * n, nys = (size_t)ncells * ns, nx = (size_t)ncells;
  if (!on_device && ncells > 0) {
    const size_t need = ny + nys + 2 * nx;
    if (need > h->stage_cap) {
      if (h->d_stage) HIPCHK(h, hipFree(h->d_stage));
      h->d_stage = nullptr;
      HIPCHK(h, hipMalloc(&h->d_stage, need * sizeof(double)));
      h->stage_cap = need;
    }
    dy = h->d_stage;
    dys = dy + ny;
    dx = dys + nys;
    dxs = dx + nx;
    HIPCHK(h, hipMemcpyAsync(dy, y, ny * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(dys, ysrc, nys * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(dx, rX, nx * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(dxs, rXsrc, nx * sizeof(double), hipMemcpyHostToDevice, s));
  }
  HIPCHK(h, hipEventRecord(h->ev0, s));
  if (ncells > 0) {
    const hipError_t e = enqueue_1d(h, dy, dys, dx, dxs, *dt_react, ncells, y, s);
    if (e != hipSuccess) return hip_fail(h, e, "react_1d launch");
  }
  HIPCHK(h, hipEventRecord(h->ev1, s));
  if (!on_device && ncells > 0) {
    HIPCHK(h, hipMemcpyAsync(y, dy, ny * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipMemcpyAsync(rX, dx, nx * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  const int r = finish(h, ncells, s, ms);
  if (h->cfg.mod_reactor) *time += *dt_react;
  return r;
}

int pp_chem_react_1d_noalloc(pp_chem* h, double* y, double* ysrc, double* rX, double* rXsrc, double* dt_react,
                             double* time, int ncells, void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_unset(h)) return r;
  if (!dt_react || !time || ncells < 0) return fail(h, PP_CHEM_EINVAL, "bad argument");
  if (ncells > 0 && (!y || !ysrc || !rX || !rXsrc)) return fail(h, PP_CHEM_EINVAL, "null array");
  if (!h->scratch) return fail(h, PP_CHEM_EHIP, "device scratch of pp_chem_react_1d_noalloc unavailable");
  const size_t counts = ((size_t)ncells * sizeof(int) + 255) & ~(size_t)255;
  if (kScratchCounts + counts > kScratchBytes)
    return fail(h, PP_CHEM_EINVAL,
                "ncells = " + std::to_string(ncells) + " exceeds the device scratch of pp_chem_react_1d_noalloc (" +
                    std::to_string((kScratchBytes - kScratchCounts) / sizeof(int)) + " cells at most)");
  int* counter = reinterpret_cast<int*>(h->scratch);
  unsigned long long* red = reinterpret_cast<unsigned long long*>(h->scratch + 256);
  int* nsteps = reinterpret_cast<int*>(h->scratch + kScratchCounts);
  unsigned char* ws = h->scratch + kScratchCounts + counts;
  const size_t ws_bytes = kScratchBytes - kScratchCounts - counts;
  // the integrator workspace in what is left: fewer resident blocks when
  // it holds fewer than the device can run
  ppc::ReactParamsDev P = h->P;
  P.ws = nullptr;
  P.ws_lanes = 0;
  const bool group = is_batched(h->cfg.mode) && h->cfg.batched_cells_per_wave != 0;
  const long long rows = group ? h->L->group_ws_rows(h->cfg.batched_cells_per_wave) : h->L->ws_rows(kmode(h->cfg.mode));
  if (rows > 0 && ncells > 0) {
    int ncu = 0;
    HIPCHK(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
    const long long per_cu =
        (is_batched(h->cfg.mode) && (group || h->cfg.batched_waves_per_simd != 1)) ? 8 : 4;
    const long long lanes_per_cell = group ? 64 / h->cfg.batched_cells_per_wave : 1;
    long long blocks = std::min<long long>(((long long)ncells * lanes_per_cell + 63) / 64, per_cu * std::max(ncu, 1));
    if (h->max_blocks > 0) blocks = std::min<long long>(blocks, h->max_blocks);
    blocks = std::min<long long>(blocks, (long long)(ws_bytes / (64 * rows * sizeof(double))));
    if (blocks < 1) return fail(h, PP_CHEM_EINVAL, "the device scratch holds no 64-lane block of this workspace");
    P.ws = reinterpret_cast<double*>(ws);
    P.ws_lanes = blocks * 64;
  }
  ppc::CellStats st{};
  st.nsteps = nsteps;  // the other per-cell counters are not kept
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(h, hipEventRecord(h->ev0, s));
  if (ncells > 0) {
    hipError_t e;
    if (is_batched(h->cfg.mode)) {
      hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, counter, 1);
      const ppc::QueueArgs q{nullptr, counter, ncells, h->max_blocks};  // cells in index order
      e = h->L->batched_1d(h->cfg.mode, P, y, ysrc, rX, rXsrc, *dt_react, q, st, s);
    } else {
      e = h->L->react_1d(kmode(h->cfg.mode), P, y, ysrc, rX, rXsrc, *dt_react, ncells, st, h->cfg.block_size, s);
    }
    if (e != hipSuccess) return hip_fail(h, e, "react_1d_noalloc launch");
  }
  HIPCHK(h, hipEventRecord(h->ev1, s));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  const int r = finish_with(h, ncells, s, ms, st, red);
  if (h->cfg.mod_reactor) *time += *dt_react;
  return r;
}

int pp_chem_flatten(pp_chem* h, const int lo[3], const int hi[3], const pp_array4* rY, const pp_array4* rYsrc,
                    const pp_array4* T, const pp_array4* rE, const pp_array4* rEsrc, double* y_vect,
                    double* src_vect, double* vect_energy, double* src_vect_energy, void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_refuses(h, "flatten")) return r;
  if (!lo || !hi || !rY || !rYsrc || !T || !rE || !rEsrc) return fail(h, PP_CHEM_EINVAL, "null argument");
  if (!rY->p || !rYsrc->p || !T->p || !rE->p || !rEsrc->p) return fail(h, PP_CHEM_EINVAL, "null array");
  ppc::FlatArgs a{};
  int ncells = 0;
  if (int r = box_args(h, lo, hi, rY, rYsrc, T, rE, rEsrc, nullptr, a.b, ncells)) return r;
  if (ncells > 0 && (!y_vect || !src_vect || !vect_energy || !src_vect_energy))
    return fail(h, PP_CHEM_EINVAL, "null flat array");
  a.y = y_vect;
  a.ysrc = src_vect;
  a.e = vect_energy;
  a.esrc = src_vect_energy;
  a.reactor_type = h->cfg.reactor_type;
  a.clean = h->cfg.clean_init_massfrac;
  hipStream_t s = (hipStream_t)stream;
  const hipError_t e = h->L->flatten(a, 0, s);
  if (e != hipSuccess) return hip_fail(h, e, "flatten launch");
  HIPCHK(h, hipStreamSynchronize(s));
  return PP_CHEM_OK;
}

int pp_chem_unflatten(pp_chem* h, const int lo[3], const int hi[3], const pp_array4* rY, const pp_array4* T,
                      const pp_array4* rE, const pp_array4* rEsrc, const pp_array4* FC, const double* y_vect,
                      const double* vect_energy, const long long* FCunt, double dt, void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_refuses(h, "unflatten")) return r;
  if (!lo || !hi || !rY || !T || !rE || !rEsrc || !FC) return fail(h, PP_CHEM_EINVAL, "null argument");
  if (!rY->p || !T->p || !rE->p || !rEsrc->p || !FC->p) return fail(h, PP_CHEM_EINVAL, "null array");
  ppc::FlatArgs a{};
  int ncells = 0;
  if (int r = box_args(h, lo, hi, rY, nullptr, T, rE, rEsrc, FC, a.b, ncells)) return r;
  if (ncells > 0 && (!y_vect || !vect_energy || !FCunt)) return fail(h, PP_CHEM_EINVAL, "null flat array");
  a.y = const_cast<double*>(y_vect);
  a.e = const_cast<double*>(vect_energy);
  a.fc = FCunt;
  a.dt = dt;
  a.reactor_type = h->cfg.reactor_type;
  a.clean = h->cfg.clean_init_massfrac;
  hipStream_t s = (hipStream_t)stream;
  const hipError_t e = h->L->flatten(a, 1, s);
  if (e != hipSuccess) return hip_fail(h, e, "unflatten launch");
  HIPCHK(h, hipStreamSynchronize(s));
  return PP_CHEM_OK;
}

int pp_chem_react_1d_chained(pp_chem* h, double* y, double* ysrc, double* rX, double* rXsrc, double* dt_react,
                             double* time, int ncells, int ncalls, void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_unset(h)) return r;
  if (!dt_react || !time || ncells < 0 || ncalls < 1) return fail(h, PP_CHEM_EINVAL, "bad argument");
  if (ncells > 0 && (!y || !ysrc || !rX || !rXsrc)) return fail(h, PP_CHEM_EINVAL, "null array");
  hipStream_t s = (hipStream_t)stream;
  if (int r = ensure_counters(h, ncells > 0 ? ncells : 1)) return r;
  if (int r = prepare_workspace(h, ncells)) return r;
  if (!h->gstream) {
    HIPCHK(h, hipStreamCreateWithFlags(&h->gstream, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreateWithFlags(&h->gev0, hipEventDisableTiming));
    HIPCHK(h, hipEventCreateWithFlags(&h->gev1, hipEventDisableTiming));
  }
  const double dt = *dt_react;
  HIPCHK(h, hipEventRecord(h->ev0, s));
  // failures are counted over every call of the chain: a cell that fails in
  // an earlier call fails the chain even if a later call integrates it
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, h->d_fail, 1);
  HIPCHK(h, hipGetLastError());
  int done = 0;
  if (ncells > 0 && is_batched(h->cfg.mode) &&
      !(h->hist_valid && h->hist_ncells == ncells && h->hist_key == y)) {
    // the first call sets the cost history the captured calls order by
    const hipError_t e = enqueue_1d(h, y, ysrc, rX, rXsrc, dt, ncells, y, s, true);
    if (e != hipSuccess) return hip_fail(h, e, "react_1d launch");
    done = 1;
  }
  if (ncells > 0 && done < ncalls) {
    const GraphKey key{{y, ysrc, rX, rXsrc}, ncells, dt, h->d_cnt, h->d_order, h->P.ws, h->P.ws_lanes};
    if (!h->gexec || !(key == h->gkey)) {
      drop_graph(h);
      HIPCHK(h, hipStreamBeginCapture(h->gstream, hipStreamCaptureModeThreadLocal));
      const hipError_t e = enqueue_1d(h, y, ysrc, rX, rXsrc, dt, ncells, y, h->gstream, true);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(h->gstream, &g);
      if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(h, e, "react_1d capture");
      }
      if (e2 != hipSuccess) return hip_fail(h, e2, "hipStreamEndCapture");
      h->graph = g;
      HIPCHK(h, hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
      h->gkey = key;
    }
    HIPCHK(h, hipEventRecord(h->gev0, s));
    HIPCHK(h, hipStreamWaitEvent(h->gstream, h->gev0, 0));
    for (int c = done; c < ncalls; ++c) HIPCHK(h, hipGraphLaunch(h->gexec, h->gstream));
    HIPCHK(h, hipEventRecord(h->gev1, h->gstream));
    HIPCHK(h, hipStreamWaitEvent(s, h->gev1, 0));
  }
  HIPCHK(h, hipEventRecord(h->ev1, s));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  int nfail = 0;
  HIPCHK(h, hipMemcpy(&nfail, h->d_fail, sizeof(int), hipMemcpyDeviceToHost));
  const int r = finish(h, ncells, s, ms);
  if (h->cfg.mod_reactor) *time += ncalls * dt;
  if (r >= 0 && nfail > 0)
    return fail(h, PP_CHEM_EINTEGRATE,
                std::to_string(nfail) + " cell integration(s) failed over the chained calls (NaN, step limit or "
                "step too small)");
  return r;
}

int pp_chem_production_rate(pp_chem* h, const double* sc, const double* T, int ncells, double* wdot, void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_refuses(h, "productionRate")) return r;
  if (ncells < 0 || (ncells > 0 && (!sc || !T || !wdot))) return fail(h, PP_CHEM_EINVAL, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  const hipError_t e = h->L->mech_eval(0, sc, T, ncells, wdot, s);
  if (e != hipSuccess) return hip_fail(h, e, "productionRate launch");
  HIPCHK(h, hipStreamSynchronize(s));
  return PP_CHEM_OK;
}

int pp_chem_progress_rate(pp_chem* h, const double* sc, const double* T, int ncells, double* q, void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_refuses(h, "the rates of progress")) return r;
  if (ncells < 0 || (ncells > 0 && (!sc || !T || !q))) return fail(h, PP_CHEM_EINVAL, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  const hipError_t e = h->L->mech_eval(3, sc, T, ncells, q, s);
  if (e != hipSuccess) return hip_fail(h, e, "progress rate launch");
  HIPCHK(h, hipStreamSynchronize(s));
  return PP_CHEM_OK;
}

int pp_chem_ajacobian(pp_chem* h, const double* sc, const double* T, int ncells, int consP, double* J,
                      void* stream) {
  if (!h) return PP_CHEM_EINVAL;
  if (int r = manifold_refuses(h, "aJacobian")) return r;
  if (ncells < 0 || (consP != 0 && consP != 1) || (ncells > 0 && (!sc || !T || !J)))
    return fail(h, PP_CHEM_EINVAL, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  const hipError_t e = h->L->mech_eval(1 + consP, sc, T, ncells, J, s);
  if (e != hipSuccess) return hip_fail(h, e, "aJacobian launch");
  HIPCHK(h, hipStreamSynchronize(s));
  return PP_CHEM_OK;
}

int pp_chem_get_stats(const pp_chem* h, pp_chem_stats* out) {
  if (!h || !out) return PP_CHEM_EINVAL;
  *out = h->stats;
  return PP_CHEM_OK;
}

int pp_chem_copy_cell_steps(pp_chem* h, int* host_out, int ncells) {
  if (!h || !host_out || ncells < 0 || ncells > h->cap) return PP_CHEM_EINVAL;
  HIPCHK(h, hipMemcpy(host_out, h->d_cnt, sizeof(int) * (size_t)ncells, hipMemcpyDeviceToHost));
  return PP_CHEM_OK;
}

const char* pp_chem_last_error(const pp_chem* h) { return h ? h->err.c_str() : "null handle"; }

}  // extern "C"
