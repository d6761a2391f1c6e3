// Host orchestration of the batched FWI path (include/fwi.h).
//
// Replaces the per-shot loop of src/main.c:396-538: every batch of shots runs
//   forward:  nt steps (one C-PML strip launch + one fused step launch each), a checkpoint of the state
//             before every K-th step, the model-multiplied terms of the last segment kept;
//   residual: res = syn - obs and the FP64 misfit;
//   adjoint:  nt steps backwards; whenever the imaging needs a forward step that is not kept, the
//             segment is recomputed from its checkpoint (K forward steps storing the imaging terms);
//   sum:      FP64 cross-shot sum of the per-shot gradients into the model gradient.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <array>
#include <vector>

#include "../../include/fwi.h"
#include "batch_kernels.h"
#include "fwi_common.h"
#include "fwi_internal.h"

#include "batch_kernels.hip"

using namespace fwi;

namespace {

constexpr int kResidualBlocks = 64;  // per shot
constexpr int kCkBlocks = 64;        // per shot, checkpoint packing
// Row pitch of a shot's fields: whole 128-byte lines, so every row starts on a line and streaming
// stores cover whole lines.
inline int row_pitch(int nx) { return (nx + 31) & ~31; }

struct Batch {
  int s0 = 0, count = 0;
  long long ncell = 0, nstrip = 0, ngather = 0, gather_off = 0;
  long long nreal = 0;          // cells of the windows (ncell counts the padded rows)
  ShotDev* d_shots = nullptr;
  int4* d_items = nullptr;      // forward work items
  int nitems = 0;
  int4* d_items_adj = nullptr;  // adjoint work items (grouped: whole workgroups per group and tile)
  int nitems_adj = 0;
  int4* d_items_core = nullptr;  // step pairs: core items (up to kPairCols output columns each)
  int nitems_core = 0;
  int4* d_items_shell = nullptr; // step pairs: single-step items of the shell, with their column ranges
  int2* d_xr_shell = nullptr;
  int nitems_shell = 0;
  int4* d_items_shell1 = nullptr;  // the same for the shell's first step, widened by the core's 4-cell layer
  int2* d_xr_shell1 = nullptr;
  int nitems_shell1 = 0;
  int4* d_items_acore = nullptr;  // adjoint step pairs: core tiles of the gradient groups (four items each)
  int nitems_acore = 0;
  int4* d_items_ash = nullptr;    // their shell: single-step items with column ranges (both steps)
  int2* d_xr_ash = nullptr;
  int nitems_ash = 0;
  long long acore_cells = 0;      // cells of the members' windows in the groups' cores
  struct Group {
    int n = 0;
    int m[4] = {0, 0, 0, 0};  // members (batch indices) in wave order; m[0] leads
    int ux0 = 0, unx = 0;     // union of the members' windows (model columns)
  };
  std::vector<Group> groups;    // gradient groups
  bool grouped = false;         // some group has more than one shot: the adjoint uses the grouped kernel
  int2* d_sblocks = nullptr;
  int nsblocks = 0;
  std::vector<ShotDev> h_shots;
  std::vector<int2> h_sblocks;
};

#define CK(x)                                          \
  do {                                                 \
    hipError_t e_ = (x);                               \
    if (e_ != hipSuccess) return fail(e_, #x);         \
  } while (0)

}  // namespace

struct fwi_plan {
  // survey
  int nx = 0, nz = 0, nt = 0, cpml = 0;
  float dh = 0, dt = 0;
  std::vector<fwi_shot> shots;
  std::vector<float> wavelet, ta, tb;
  fwi_options opt{};
  int K = 1, J = 1;  // checkpoint interval and number of segments
  int S = 1;         // trailing segments whose imaging terms the forward keeps (no checkpoint, no recomputation)
  // with s trailing segments kept (default: the plan's S)
  int sfirst(int s) const { return (J - s) * K; }                  // first step kept from the forward
  int qslots(int s) const { return std::max(nt - sfirst(s), K); }  // imaging-term slots (2 fields each)
  int nckpt(int s) const { return std::max(J - s - 1, 0); }        // checkpoints 1..J-s-1
  int sfirst() const { return sfirst(S); }
  int qslots() const { return qslots(S); }
  int nckpt() const { return nckpt(S); }
  int tz = 32;       // rows per work item (forward kernels)
  int order = 0;     // work-item order (build_items)
  int tz_adj = 32;   // rows per work item and order of the adjoint kernels
  int order_adj = 0;
  int tz_pair = 32;  // rows per fused core item (step pairs)
  int tz_shell = 0;  // rows per shell item of the launch completing a step pair (0: tz)
  bool pairs_off = false;  // step_pairs = 1 and the tuner timed single steps faster (small, latency-bound grids)
  int tz_apair = 32;       // rows per core item of the adjoint step pairs
  bool apairs_off = false; // adj_step_pairs = 1 and the tuner timed single adjoint steps faster
  // FP64 near-surface band of the increment form: rows [4, band) of every shot (0: none).  State: seven FP64 arrays
  // of cap_band cells (band rows x row pitch of every shot): P, F and psi_x by step parity, zeta_x.
  int band = 0;
  long long cap_band = 0;
  double* d_band = nullptr;
  double* bP(int n) const { return d_band + (size_t)(n & 1) * cap_band; }
  double* bF(int n) const { return d_band + (size_t)(2 + (n & 1)) * cap_band; }
  double* bpsx(int n) const { return d_band + (size_t)(4 + (n & 1)) * cap_band; }
  double* bzx() const { return d_band + 6 * (size_t)cap_band; }
  static constexpr int kBandF = 7;
  bool tuned = false;
  std::vector<Batch> batches;
  long long gather_total = 0;
  // device
  float* d_src = nullptr;
  float* d_ta = nullptr;
  float* d_tb = nullptr;
  float* d_work = nullptr;  // all per-cell workspace of the largest batch
  size_t work_bytes = 0;
  long long cap_cell = 0, cap_strip = 0, cap_gather = 0;
  double* d_partial = nullptr;   // kResidualBlocks per shot of the largest batch
  double* d_misfit = nullptr;
  double* d_shot_misfit = nullptr;  // per shot of the survey, from the last misfit or gradient
  size_t ck_esize = 4;              // bytes per checkpoint value (4 FP32, 2 BF16 / FP16)
  int max_count = 0;                // shots in the largest batch
  unsigned* d_ck_amax = nullptr;    // 16-bit checkpoints: max-abs bits per (checkpoint, shot, field)
  // sampled kernel timing: kinds 0 = forward step (no store), 1 = forward step (store / recompute), each one launch
  // with the FP64 near-surface band, 2 = adjoint step, 3 / 4 = step pair launch without / with store (the core's two
  // steps, the band's and the shell's first), 6 = the two adjoint C-PML strip kernels, 8 / 9 = the launch that
  // completes a step pair without / with store (the band's and the shell's second step); every `timing_stride`-th
  // launch of a kind is bracketed by events on its stream.  Kinds 5 and 7 (separate C-PML strip and band launches of
  // earlier versions) are unused.
  static constexpr int kKinds = 10;
  int timing_stride = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_open;  // (kind, event index of the start; end = start + 1)
  double kt_ms[kKinds] = {};
  long long kt_count[kKinds] = {};
  long long launch_no[kKinds] = {};

  int timed_begin(int kind, hipStream_t st) {
    if (timing_stride <= 0 || (launch_no[kind]++ % timing_stride) != 0) return -1;
    size_t idx = ev_open.size() * 2;
    while (ev_pool.size() < idx + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      ev_pool.push_back(e);
    }
    (void)hipEventRecord(ev_pool[idx], st);
    ev_open.push_back({kind, (int)idx});
    return (int)idx;
  }
  void timed_end(int idx, hipStream_t st) {
    if (idx >= 0) (void)hipEventRecord(ev_pool[idx + 1], st);
  }
  void timing_collect() {
    if (ev_open.empty()) return;
    (void)hipEventSynchronize(ev_pool[ev_open.back().second + 1]);
    for (auto& o : ev_open) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev_pool[o.second], ev_pool[o.second + 1]) == hipSuccess) {
        kt_ms[o.first] += ms;
        kt_count[o.first] += 1;
      }
    }
    ev_open.clear();
  }

  // workspace views (sized for cap_cell)
  // forward fields: the state of step n in set n mod 3 in the increment form (P 0..2, F 3..5, Q = P^n - P^{n-1}
  // 6..8, R = F^n - F^{n-1} 9..11; three sets so that step pairs can keep step n while writing n+1 and n+2), in set
  // n mod 2 of P 0..1 and F 3..4 in the two-level form; C-PML memory psi_x^n 12..13 and psi_z^n 14..15 by step
  // parity (work items compute psi^{n+1} from psi^n, their neighbours' included), zx 16, zz 17
  // adjoint fields of the exact adjoint: mu_P^j in 0..3 and mu_F^j in 4..7, set (j+1) mod 4 (four time levels, so
  // that a step pair reads two and writes two others); the C-PML memory variables psi-bar_x 8..9, zeta-bar_x
  // 10..11, psi-bar_z 12..13, zeta-bar_z 14..15 by step parity; the increment form's D_P 16, D_F 17.  The
  // reference-compatible adjoint: mu_P 0..1, mu_F 2..3 by step parity, its eight C-PML fields in 4..11.
  static constexpr int kFwdF = 18, kAdjF = 18, kGradF = 3, kFixedF = kFwdF + kAdjF + kGradF;
  static int aset(int j) { return (j + 1) & 3; }
  int fset(int n) const { return inc() ? n % 3 : (n & 1); }
  float* fP(int n) const { return fwd(fset(n)); }
  float* fF(int n) const { return fwd(3 + fset(n)); }
  float* fQ(int n) const { return fwd(6 + n % 3); }
  float* fR(int n) const { return fwd(9 + n % 3); }
  float* fpsx(int n) const { return fwd(12 + (n & 1)); }
  float* fpsz(int n) const { return fwd(14 + (n & 1)); }
  float* fzx() const { return fwd(16); }
  float* fzz() const { return fwd(17); }
  float* fwd(int k) const { return d_work + (size_t)k * cap_cell; }
  float* adj(int k) const { return d_work + (size_t)(kFwdF + k) * cap_cell; }                // 0..17
  float* sgrad(int k) const { return d_work + (size_t)(kFwdF + kAdjF + k) * cap_cell; }      // 0..2
  float* qbuf() const { return d_work + (size_t)kFixedF * cap_cell; }                      // qslots * 2 * cap_cell
  // checkpoints j = 1..J-1: 4 cell fields then 4 strip fields, FP32 or 16-bit (ck_esize bytes per value)
  size_t ck_bytes() const {
    return (4 * (size_t)cap_cell + 4 * (size_t)cap_strip) * ck_esize + kBandF * (size_t)cap_band * sizeof(double);
  }
  char* ckpt(int j) const {
    return (char*)(d_work + (size_t)(kFixedF + 2 * (size_t)qslots()) * cap_cell) + (size_t)(j - 1) * ck_bytes();
  }
  float* gathers() const {
    return (float*)((char*)(d_work + (size_t)(kFixedF + 2 * (size_t)qslots()) * cap_cell) + (size_t)nckpt() * ck_bytes());
  }
  size_t workspace_floats(int s) const {
    return (size_t)(kFixedF + 2 * (size_t)qslots(s)) * cap_cell + (ck_bytes() * (size_t)nckpt(s) + 3) / 4 +
           2 * (size_t)cap_gather;
  }
  float* resid() const { return gathers() + cap_gather; }
  bool inc() const { return opt.arith == FWI_ARITH_FP32I; }
  // the four cell fields a checkpoint holds before step n: both levels of P and F, or in the increment
  // form the set of step n (P, F and their increments Q, R)
  int ckpt_field(int f, int n) const {
    if (!inc()) {
      const int map[4] = {0, 1, 3, 4};
      return map[f];
    }
    const int m = n % 3;
    const int map[4] = {m, 3 + m, 6 + m, 9 + m};
    return map[f];
  }
  std::string err;

  int fail(hipError_t e, const char* what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    set_message(err);
    return (int)e;
  }
  int fail(const std::string& msg) {
    err = msg;
    set_message(err);
    return 1;
  }

  int zero(float* p, size_t n, hipStream_t st) {
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    if (n) hipLaunchKernelGGL(k_zero_f32, dim3(g), dim3(256), 0, st, p, (long long)n);
    CK(hipGetLastError());
    return 0;
  }
  int copy(float* d, const float* src, size_t n, hipStream_t st) {
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    if (n) hipLaunchKernelGGL(k_copy_f32, dim3(g), dim3(256), 0, st, d, src, (long long)n);
    CK(hipGetLastError());
    return 0;
  }

  // ---------------------------------------------------------------------------------------------
  // Graphs: the device work of one batch of one entry point is captured once, on the plan's own stream, and
  // its executable graphs are cached under the entry point, the batch and the caller's buffers; later calls
  // with the same buffers (an inversion updates its model in place) replay them on the caller's stream.  A
  // capture is cut into consecutive graphs of at most kGraphSteps time steps (forward, recomputation or adjoint;
  // 2-4 kernels each), replayed in order, so no graph grows with nt: a 4000-step survey1600 gradient is three
  // graphs of ~12 000 kernel nodes instead of one of ~35 000.  Not used while kernel timing is on.
  struct GraphEntry {
    std::array<const void*, 10> key;
    std::vector<hipGraphExec_t> execs;
  };
  static constexpr size_t kMaxGraphs = 8;
  static constexpr int kGraphSteps = 4096;
  std::vector<GraphEntry> graphs;
  hipStream_t cap_stream = nullptr;
  long long graph_captures = 0, graph_replays = 0;
  long long graph_nodes_largest = 0, graph_nodes_last = 0;  // kernel nodes: largest graph, last capture's total
  bool capturing = false;
  int cap_steps = 0;                   // time steps in the graph being captured
  std::vector<hipGraphExec_t> cap_execs;  // graphs of the capture so far

  int end_capture() {
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(cap_stream, &graph);
    if (e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(e, "hipStreamEndCapture");
    }
    size_t nodes = 0;
    (void)hipGraphGetNodes(graph, nullptr, &nodes);
    graph_nodes_largest = std::max(graph_nodes_largest, (long long)nodes);
    graph_nodes_last += (long long)nodes;
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(ei, "hipGraphInstantiate");
    cap_execs.push_back(exec);
    return 0;
  }
  // Called between time steps: while capturing, a graph that holds kGraphSteps steps is closed and the capture
  // continues into the next one on the same stream (the work is one stream's sequence, so cutting it anywhere
  // between launches keeps its order).
  int step_done() {
    if (!capturing || ++cap_steps < kGraphSteps) return 0;
    if (int r = end_capture()) return r;
    CK(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
    cap_steps = 0;
    return 0;
  }

  template <class F>
  int run_graph(const std::array<const void*, 10>& key, hipStream_t st, F&& body) {
    if (!opt.use_graph || timing_stride > 0) return body(st);
    for (size_t i = 0; i < graphs.size(); ++i)
      if (graphs[i].key == key) {
        std::rotate(graphs.begin() + i, graphs.begin() + i + 1, graphs.end());  // most recently used last
        for (hipGraphExec_t x : graphs.back().execs) CK(hipGraphLaunch(x, st));
        ++graph_replays;
        return 0;
      }
    if (!cap_stream) CK(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
    cap_execs.clear();
    graph_nodes_last = 0;
    cap_steps = 0;
    CK(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
    capturing = true;
    int r = body(cap_stream);
    capturing = false;
    if (r) {
      hipGraph_t graph = nullptr;
      if (hipStreamEndCapture(cap_stream, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
    } else {
      r = end_capture();
    }
    if (r) {
      for (hipGraphExec_t x : cap_execs) (void)hipGraphExecDestroy(x);
      cap_execs.clear();
      return r;
    }
    if (graphs.size() >= kMaxGraphs) {  // least recently used first
      // its last launch may still run on the stream it was given: wait for the device before destroying it
      (void)hipDeviceSynchronize();
      for (hipGraphExec_t x : graphs.front().execs) (void)hipGraphExecDestroy(x);
      graphs.erase(graphs.begin());
    }
    graphs.push_back({key, cap_execs});
    cap_execs.clear();
    ++graph_captures;
    for (hipGraphExec_t x : graphs.back().execs) CK(hipGraphLaunch(x, st));
    return 0;
  }

  ~fwi_plan() {
    (void)hipDeviceSynchronize();  // launches of this plan may still be running on the caller's streams
    for (auto& g : graphs)
      for (hipGraphExec_t x : g.execs) (void)hipGraphExecDestroy(x);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (auto& b : batches) {
      (void)hipFree(b.d_shots);
      (void)hipFree(b.d_items);
      (void)hipFree(b.d_items_adj);
      (void)hipFree(b.d_items_core);
      (void)hipFree(b.d_items_shell);
      (void)hipFree(b.d_xr_shell);
      (void)hipFree(b.d_items_shell1);
      (void)hipFree(b.d_xr_shell1);
      (void)hipFree(b.d_items_acore);
      (void)hipFree(b.d_items_ash);
      (void)hipFree(b.d_xr_ash);
      (void)hipFree(b.d_sblocks);
    }
    (void)hipFree(d_src);
    (void)hipFree(d_ta);
    (void)hipFree(d_tb);
    (void)hipFree(d_work);
    (void)hipFree(d_partial);
    (void)hipFree(d_misfit);
    (void)hipFree(d_shot_misfit);
    (void)hipFree(d_ck_amax);
    (void)hipFree(d_band);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
  }

  // ---------------------------------------------------------------------------------------------
  // Work items: (shot, 64-column band, first row, end row) per wave, four per workgroup.
  // Row runs of `rows` rows over [kRim, nz-kRim); a run is cut at every split row strictly inside it.
  static std::vector<std::pair<int, int>> row_runs(int nz, int rows, std::vector<int> splits, int zfirst = kRim) {
    std::sort(splits.begin(), splits.end());
    std::vector<std::pair<int, int>> runs;
    for (int z0 = zfirst; z0 < nz - kRim; z0 += rows) {
      const int z1 = std::min(z0 + rows, nz - kRim);
      int a = z0;
      for (int sp : splits)
        if (sp > a && sp < z1) {
          runs.push_back({a, sp});
          a = sp;
        }
      runs.push_back({a, z1});
    }
    return runs;
  }

  static std::vector<int4> ungrouped_items(const Batch& b, int nz, int rows, int order,
                                           const std::vector<int>& zsplit = {}, int zfirst = kRim) {
    std::vector<int4> items;
    int maxnx = 0;
    for (int s = 0; s < b.count; ++s) maxnx = std::max(maxnx, b.h_shots[s].nx);
    const std::vector<std::pair<int, int>> runs = row_runs(nz, rows, zsplit, zfirst);
    if (order == 1) {
      for (int s = 0; s < b.count; ++s)
        for (auto& r : runs)
          for (int x0 = 0; x0 < b.h_shots[s].nx; x0 += 64) items.push_back(make_int4(s, x0, r.first, r.second));
    } else if (order == 3 || order == 4) {
      for (int g = 0; g < b.count; g += 4) {
        const int ge = std::min(b.count, g + 4);
        int gnx = 0;
        for (int s = g; s < ge; ++s) gnx = std::max(gnx, b.h_shots[s].nx);
        for (auto& r : runs)
          for (int x0 = 0; x0 < gnx; x0 += 64)
            for (int s = g; s < ge; ++s)
              if (x0 < b.h_shots[s].nx) items.push_back(make_int4(s, x0, r.first, r.second));
      }
    } else if (order == 2) {
      for (int x0 = 0; x0 < maxnx; x0 += 64) {
        std::vector<int> sh;
        for (int s = 0; s < b.count; ++s)
          if (x0 < b.h_shots[s].nx) sh.push_back(s);
        for (size_t g = 0; g < sh.size(); g += 4)
          for (auto& r : runs)
            for (size_t k = g; k < std::min(sh.size(), g + 4); ++k) items.push_back(make_int4(sh[k], x0, r.first, r.second));
      }
    } else {
      for (auto& r : runs)
        for (int x0 = 0; x0 < maxnx; x0 += 64)
          for (int s = 0; s < b.count; ++s)
            if (x0 < b.h_shots[s].nx) items.push_back(make_int4(s, x0, r.first, r.second));
    }
    return items;
  }

  // Grouped adjoint items: every workgroup is one gradient group on one tile, member k on wave k, and
  // -(leader+1) (a placeholder wave) where the group has fewer than 4 shots.  Orders: 0 row run -> band ->
  // group; 2 band -> group -> row run; 3 group -> row run -> band.
  static std::vector<int4> grouped_items(const Batch& b, int nz, int rows, int order,
                                         const std::vector<int>& zsplit) {
    std::vector<int4> items;
    const std::vector<std::pair<int, int>> runs = row_runs(nz, rows, zsplit);
    int maxnx = 0;
    for (int s = 0; s < b.count; ++s) maxnx = std::max(maxnx, b.h_shots[s].nx);
    // one workgroup: the group's members on the 64 model columns from bm, each at its window's offset
    auto push = [&](const Batch::Group& g, int bm, const std::pair<int, int>& r) {
      for (int k = 0; k < 4; ++k) {
        bool member = k < g.n;
        if (member) {
          const ShotDev& d = b.h_shots[g.m[k]];
          member = bm < d.x0 + d.nx && bm + 64 > d.x0;
        }
        const int s = member ? g.m[k] : g.m[0];
        items.push_back(make_int4(member ? s : -(s + 1), bm - b.h_shots[s].x0, r.first, r.second));
      }
    };
    int maxunx = 0;
    for (auto& g : b.groups) maxunx = std::max(maxunx, g.unx);
    if (order == 2) {
      for (int x = 0; x < maxunx; x += 64)
        for (auto& g : b.groups)
          if (x < g.unx)
            for (auto& r : runs) push(g, g.ux0 + x, r);
    } else if (order == 3 || order == 4) {
      for (auto& g : b.groups)
        for (auto& r : runs)
          for (int x = 0; x < g.unx; x += 64) push(g, g.ux0 + x, r);
    } else {
      for (auto& r : runs)
        for (int x = 0; x < maxunx; x += 64)
          for (auto& g : b.groups)
            if (x < g.unx) push(g, g.ux0 + x, r);
    }
    (void)maxnx;
    return items;
  }

  // Work items with C-PML cells take longer per row than interior items.  Workgroups are dealt to the XCDs in
  // contiguous ranges of the item order (xcd_logical_block) and each XCD dispatches its range in order, so
  // within every range the units (of `unit` items: 4 for a workgroup that must stay together) containing such
  // an item are moved to the front, keeping the order otherwise; their longer rows then overlap the rest.
  template <class IsEdge>
  static void edge_first(std::vector<int4>& items, int unit, IsEdge&& is_edge) {
    const long long nwg = ((long long)items.size() + 3) / 4;
    const long long q = nwg / 8, r = nwg % 8;
    std::vector<int4> out;
    out.reserve(items.size());
    for (long long x = 0; x < 8; ++x) {
      const long long w0 = x * q + std::min(x, r), w1 = w0 + q + (x < r ? 1 : 0);
      const size_t i0 = (size_t)std::min<long long>(4 * w0, (long long)items.size());
      const size_t i1 = (size_t)std::min<long long>(4 * w1, (long long)items.size());
      std::vector<int4> rest;
      for (size_t u = i0; u < i1; u += unit) {
        const size_t e = std::min(i1, u + unit);
        bool edge = false;
        for (size_t k = u; k < e; ++k) edge = edge || is_edge(items[k]);
        (edge ? out : rest).insert((edge ? out : rest).end(), items.begin() + u, items.begin() + e);
      }
      out.insert(out.end(), rest.begin(), rest.end());
    }
    items.swap(out);
  }

  int upload_items(int4*& dst, int& n, const std::vector<int4>& items) {
    (void)hipFree(dst);
    dst = nullptr;
    n = (int)items.size();
    if (items.empty()) return 0;
    CK(hipMalloc(&dst, items.size() * sizeof(int4)));
    CK(hipMemcpy(dst, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
    return 0;
  }

  // Work items of the forward kernels in one of five orders (the tuner picks one per plan):
  //   0  row run -> band -> shot: the four waves of a workgroup are four shots on one tile (model rows
  //      shared through the L1), x-neighbour tiles are a few workgroups apart;
  //   1  shot -> row run -> band: a workgroup is four adjacent bands of one shot (x halos shared through
  //      the L1), z-neighbour row runs one band row apart;
  //   2  band -> shot group of 4 -> row run -> shot: consecutive workgroups walk down a column of tiles,
  //      so z halos are reused from the L2 while the neighbour is still resident;
  //   3  shot group of 4 -> row run -> band -> shot: a workgroup is four shots on one tile (model rows shared
  //      through the L1) and consecutive workgroups are its x neighbours, with the z neighbours a row of bands
  //      later, so both halos are read while the neighbour tile is resident in the L2;
  //   4  as 3 with every workgroup one tile of four shots with the same window: the four waves of an interior
  //      tile share the model rows through LDS (kShareModel).  Offered when the batch's shots, ordered by window,
  //      form such groups of four (full-width surveys).
  // Workgroup ids are remapped so each XCD receives a contiguous range of this order.
  // Forward items, cut at row nz-e (e = max(cpml,4), the first bottom C-PML row): the kernel runs items without
  // C-PML cells with its pipelined interior body; the others go first in each XCD's range.
  // Shot groups of order 4: the batch's shots ordered by window (x0, width, then index) in runs of four with
  // identical windows; empty if the batch does not split that way.
  static std::vector<std::array<int, 4>> share_groups(const Batch& b) {
    std::vector<int> ord(b.count);
    for (int k = 0; k < b.count; ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](int u, int v) {
      const ShotDev &x = b.h_shots[u], &y = b.h_shots[v];
      return x.x0 != y.x0 ? x.x0 < y.x0 : x.nx < y.nx;
    });
    std::vector<std::array<int, 4>> g;
    if (b.count % 4) return g;
    for (int k = 0; k < b.count; k += 4) {
      for (int j = 1; j < 4; ++j)
        if (b.h_shots[ord[k + j]].x0 != b.h_shots[ord[k]].x0 || b.h_shots[ord[k + j]].nx != b.h_shots[ord[k]].nx)
          return {};
      g.push_back({ord[k], ord[k + 1], ord[k + 2], ord[k + 3]});
    }
    return g;
  }
  int build_items(Batch& b, int rows, int order) {
    const int e = std::max(cpml, kRim);
    auto interior = [&](const int4& it) {
      return it.y >= e && it.y + 64 <= b.h_shots[it.x & ~kShareModel].nx - e && it.w <= nz - e;
    };
    std::vector<int4> items;
    const auto groups = order == 4 ? share_groups(b) : std::vector<std::array<int, 4>>{};
    const int zfirst = std::max(kRim, band);  // rows above are the FP64 band's
    if (!groups.empty()) {
      const std::vector<std::pair<int, int>> runs = row_runs(nz, rows, {nz - e}, zfirst);
      for (const auto& g : groups)
        for (auto& r : runs)
          for (int x0 = 0; x0 < b.h_shots[g[0]].nx; x0 += 64) {
            const bool sh = interior(make_int4(g[0], x0, r.first, r.second));
            for (int k = 0; k < 4; ++k) items.push_back(make_int4(g[k] | (sh ? kShareModel : 0), x0, r.first, r.second));
          }
      edge_first(items, 4, [&](const int4& it) { return !interior(it); });
    } else {
      items = ungrouped_items(b, nz, rows, order == 4 ? 3 : order, {nz - e}, zfirst);
      edge_first(items, 1, [&](const int4& it) { return !interior(it); });
    }
    return upload_items(b.d_items, b.nitems, items);
  }
  // Step pairs.  The core of a window is columns [pair_cx0, pair_cx1) and rows [pair_za(), pair_zb()); the shell,
  // everything else of the float state, takes single steps.  Core items: (shot, first of up to kPairCols output
  // columns, row run of `rows`), shot -> row run -> tile, one workgroup each.  Shell items: 64-column bands with the
  // column range each one updates, over the rows above and below the core and the side columns beside it (a window
  // without a core is all shell).  The shell's first step also covers the core's cells within 4 of the shell, whose
  // step n+1 values the shell's second step and the C-PML update read.  The first step's shell items are runs of
  // `rows` rows, as the core's: they share the launch with the bandwidth-bound core workgroups, and fewer, longer
  // items spend less of it on their queue warm-up; the second step's are shorter runs (tz_shell rows, else tz):
  // alone in their launch, more items finish it sooner.
  int pair_za() const { return std::max(kRim, band) + 4; }
  int pair_zb() const { return nz - cpml - 4; }
  int build_pair_items(Batch& b, int rows) {
    std::vector<int4> core, shell[2];
    std::vector<int2> xr[2];
    const int zf = std::max(kRim, band), za = pair_za(), zb = pair_zb();
    auto runs_of = [&](int z0, int z1, int step) {
      std::vector<std::pair<int, int>> r;
      for (int z = z0; z < z1; z += step) r.push_back({z, std::min(z + step, z1)});
      return r;
    };
    const auto core_runs = runs_of(za, zb, rows);
    for (int s = 0; s < b.count; ++s) {
      const int nxs = b.h_shots[s].nx, cx0 = pair_cx0(cpml), cx1 = pair_cx1(nxs, cpml);
      auto band_items = [&](int k, int x, int x1, int z0, int z1, int lo, int hi) {
        for (; x < x1; x += 64)
          for (auto& r : runs_of(z0, z1, k == 0 ? rows : tz_shell > 0 ? tz_shell : tz)) {
            shell[k].push_back(make_int4(s, x, r.first, r.second));
            xr[k].push_back(make_int2(lo, hi));
          }
      };
      if (cx1 <= cx0 || zb <= za) {  // no core: the whole window is shell
        for (int k = 0; k < 2; ++k) band_items(k, 0, nxs, zf, nz - kRim, 0, nxs);
        continue;
      }
      for (auto& r : core_runs)
        for (int x = cx0; x < cx1; x += kPairCols) core.push_back(make_int4(s, x, r.first, r.second));
      for (int k = 0; k < 2; ++k) {
        // The first step's shell reaches 4 cells into the core.  Its four parts must not overlap (a C-PML memory
        // variable updated twice in one launch would be wrong): in a core thinner than 8 rows or columns the part
        // above (left of) the core takes the rows (columns) both would reach.
        const int w = k == 0 ? 4 : 0;
        const int z1 = std::min(za + w, zb), z2 = std::max(zb - w, z1);
        const int x1 = std::min(cx0 + w, cx1), x2 = std::max(cx1 - w, x1);
        band_items(k, 0, nxs, zf, z1, 0, nxs);                      // above the core
        band_items(k, 0, nxs, z2, nz - kRim, 0, nxs);               // below it
        band_items(k, 0, ((x1 + 63) / 64) * 64, z1, z2, 0, x1);     // left of it
        band_items(k, (x2 / 64) * 64, nxs, z1, z2, x2, nxs);        // right of it
      }
    }
    if (int r = upload_items(b.d_items_core, b.nitems_core, core)) return r;
    if (int r = upload_items(b.d_items_shell1, b.nitems_shell1, shell[0])) return r;
    if (int r = upload_items(b.d_items_shell, b.nitems_shell, shell[1])) return r;
    int2** dst[2] = {&b.d_xr_shell1, &b.d_xr_shell};
    for (int k = 0; k < 2; ++k) {
      (void)hipFree(*dst[k]);
      *dst[k] = nullptr;
      if (!xr[k].empty()) {
        CK(hipMalloc(dst[k], xr[k].size() * sizeof(int2)));
        CK(hipMemcpy(*dst[k], xr[k].data(), xr[k].size() * sizeof(int2), hipMemcpyHostToDevice));
      }
    }
    return 0;
  }

  // Work items of the adjoint kernels: grouped (orders 0, 2 and 3) when the batch has gradient groups.  No item
  // crosses row nz-cpml-4, the first row whose z stencil reaches the bottom C-PML strip, nor the end row of the
  // FP64 band: the exact adjoint kernel runs the rows of an item with or without the bottom C-PML terms, by its
  // first row, and in the increment form when it lies in the band.
  int build_adj_items(Batch& b, int rows, int order) {
    const int zsplit = nz - cpml - 4;
    std::vector<int> splits = {zsplit};
    if (band > kRim) splits.push_back(band);  // the band's rows take the increment form
    std::vector<int4> items = b.grouped ? grouped_items(b, nz, rows, order, splits)
                                        : ungrouped_items(b, nz, rows, order, splits);
    if (opt.grad_mode == FWI_GRAD_EXACT)
      edge_first(items, b.grouped ? 4 : 1, [&](const int4& it) {
        if (it.x < 0) return false;  // a placeholder wave computes nothing
        const int nxs = b.h_shots[it.x].nx;
        return it.y < cpml + 4 || it.y + 64 > nxs - cpml - 4 || it.z >= zsplit;
      });
    return upload_items(b.d_items_adj, b.nitems_adj, items);
  }

  // Adjoint step pairs (grouped batches).  A group's core is the model columns where every member's window has
  // [adj_pair_cx0, adj_pair_cx1) and rows [apair_za(), apair_zb()); the core's work items store mu^k and mu^{k-1}
  // on all of its cells, so the shell -- everything else of the group's union window -- takes single steps k and
  // k-1 over the same items (64-column bands of the union with the column range each updates, four members per
  // workgroup as the single-step items, cut at the band and at row nz-cpml-4).  Core items: one tile of up to
  // kAdjPairCols columns and `rows` rows per workgroup, group -> row run -> tile.  The partition is the same for
  // every member, so each cell's four imaging terms are still added in one place and order.
  int apair_za() const { return std::max(kRim, band) + 4; }
  int apair_zb() const { return nz - cpml - 8; }
  int build_adj_pair_items(Batch& b, int rows) {
    std::vector<int4> core, shell;
    std::vector<int2> xr;
    const int za = apair_za(), zb = apair_zb(), zc = nz - cpml - 4;
    b.acore_cells = 0;
    auto runs_in = [&](int z0, int z1, int step) {  // runs of `step` rows in [z0, z1), cut at zc and the band end
      std::vector<std::pair<int, int>> r;
      for (int z = z0; z < z1; z += step) {
        int a = z;
        const int e = std::min(z + step, z1);
        for (int sp : {band > kRim ? band : -1, zc})
          if (sp > a && sp < e) {
            r.push_back({a, sp});
            a = sp;
          }
        r.push_back({a, e});
      }
      return r;
    };
    for (const auto& g : b.groups) {
      const int gx0 = g.ux0, gx1 = g.ux0 + g.unx;
      int c0 = gx0, c1 = gx1;
      for (int j = 0; j < g.n; ++j) {
        const ShotDev& d = b.h_shots[g.m[j]];
        c0 = std::max(c0, d.x0 + adj_pair_cx0(cpml));
        c1 = std::min(c1, d.x0 + adj_pair_cx1(d.nx, cpml));
      }
      const bool has_core = c1 > c0 && zb > za;
      // the group's members (placeholder waves where one is missing or its window misses the band) on model
      // columns [bm, bm+64), rows [z0, z1), updating model columns [lo, hi)
      auto band_items = [&](int bm, int z0, int z1, int lo, int hi) {
        for (auto& r : runs_in(z0, z1, tz_adj))
          for (int j = 0; j < 4; ++j) {
            bool member = j < g.n;
            if (member) {
              const ShotDev& d = b.h_shots[g.m[j]];
              member = bm < d.x0 + d.nx && bm + 64 > d.x0;
            }
            const int s = member ? g.m[j] : g.m[0];
            const int xs = b.h_shots[s].x0;
            shell.push_back(make_int4(member ? s : -(s + 1), bm - xs, r.first, r.second));
            xr.push_back(make_int2(lo - xs, hi - xs));
          }
      };
      auto region = [&](int x0, int x1, int z0, int z1) {  // the bands meeting model columns [x0, x1)
        if (x1 <= x0 || z1 <= z0) return;
        for (int bm = gx0 + ((x0 - gx0) / 64) * 64; bm < x1; bm += 64) band_items(bm, z0, z1, x0, x1);
      };
      if (!has_core) {
        region(gx0, gx1, kRim, nz - kRim);
        continue;
      }
      region(gx0, gx1, kRim, za);       // above the core
      region(gx0, gx1, zb, nz - kRim);  // below it
      region(gx0, c0, za, zb);          // left of it
      region(c1, gx1, za, zb);          // right of it
      b.acore_cells += (long long)(c1 - c0) * (zb - za) * g.n;
      for (int z = za; z < zb; z += rows)
        for (int x = c0; x < c1; x += kAdjPairCols) {
          const int ncols = std::min(kAdjPairCols, c1 - x);
          for (int j = 0; j < 4; ++j) {
            const int s = j < g.n ? g.m[j] : g.m[0];
            const int xs = b.h_shots[s].x0;
            core.push_back(make_int4(j < g.n ? s : -(s + 1), (x - xs) | (ncols << 16), z, std::min(z + rows, zb)));
          }
        }
    }
    if (int r = upload_items(b.d_items_acore, b.nitems_acore, core)) return r;
    if (int r = upload_items(b.d_items_ash, b.nitems_ash, shell)) return r;
    (void)hipFree(b.d_xr_ash);
    b.d_xr_ash = nullptr;
    if (!xr.empty()) {
      CK(hipMalloc(&b.d_xr_ash, xr.size() * sizeof(int2)));
      CK(hipMemcpy(b.d_xr_ash, xr.data(), xr.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    return 0;
  }

  GridArgs grid_args(const Batch& b) const {
    GridArgs g;
    g.shots = b.d_shots;
    g.nz = nz;
    g.cpml = cpml;
    g.dh = dh;
    g.dh2 = dh * dh;
    g.ta = d_ta;
    g.tb = d_tb;
    return g;
  }

  // ---------------------------------------------------------------------------------------------
  // Arguments of forward step n over an item list; with_band: the FP64 band's step n in the same launch.
  FwdArgs fwd_args(const Batch& b, const ModelArgs& m, int n, float* qa, float* qb, float* gather, const int4* items,
                   int nitems, const int2* xr, bool with_band) const {
    FwdArgs a;
    a.g = grid_args(b);
    a.m = m;
    a.items = items;
    a.nitems = nitems;
    a.C = (dt * dt) / (dh * dh);
    a.src = d_src;
    a.Pc = fP(n);
    a.Po = fP(n + 1);
    a.Fc = fF(n);
    a.Fo = fF(n + 1);
    a.psx = fpsx(n);
    a.psz = fpsz(n);
    a.psxn = fpsx(n + 1);
    a.pszn = fpsz(n + 1);
    a.zx = fzx();
    a.zz = fzz();
    a.Qc = fQ(n);
    a.Rc = fR(n);
    a.Qo = fQ(n + 1);
    a.Ro = fR(n + 1);
    a.gather = gather;
    a.qa = qa;
    a.qb = qb;
    a.it = n;
    a.xr = xr;
    a.nband_wg = a.nband_real = a.nbc = 0;
    if (with_band && band > kRim) {
      a.band = band_args(b, m, n, qa, qb, gather);
      int maxnx = 0;
      for (const ShotDev& d : b.h_shots) maxnx = std::max(maxnx, d.nx);
      a.nbc = (maxnx + 255) / 256;
      a.nband_real = a.nbc * band_groups() * b.count;
      a.nband_wg = (a.nband_real + 7) / 8 * 8;
    }
    return a;
  }
  // one forward launch: the FP64 band's step n and the items
  int fwd_main(const Batch& b, const ModelArgs& m, int n, int qm, float* qa, float* qb, float* gather,
               const int4* items, int nitems, const int2* xr, int kind, hipStream_t st) {
    const FwdArgs a = fwd_args(b, m, n, qa, qb, gather, items, nitems, xr, true);
    if (nitems == 0 && a.nband_wg == 0) return 0;
    dim3 grid((unsigned)(a.nband_wg + (nitems + 3) / 4));
    const int am = opt.arith;
    const int tev = kind >= 0 ? timed_begin(kind, st) : -1;
#define FWD_LAUNCH(QMV, SV) hipLaunchKernelGGL((k_fwd_main<QMV, SV>), grid, dim3(256), 0, st, a)
#define FWD_LAUNCH_INC(QMV) hipLaunchKernelGGL((k_fwd_main<QMV, float, true>), grid, dim3(256), 0, st, a)
    if (am == FWI_ARITH_FP32I) {
      if (qm == 0) FWD_LAUNCH_INC(0);
      else if (qm == 1) FWD_LAUNCH_INC(1);
      else FWD_LAUNCH_INC(2);
    } else if (am == FWI_ARITH_MIXED) {
      if (qm == 0) FWD_LAUNCH(0, double);
      else if (qm == 1) FWD_LAUNCH(1, double);
      else FWD_LAUNCH(2, double);
    } else if (am == FWI_ARITH_FP32D) {
      if (qm == 0) FWD_LAUNCH(0, F32D);
      else if (qm == 1) FWD_LAUNCH(1, F32D);
      else FWD_LAUNCH(2, F32D);
    } else {
      if (qm == 0) FWD_LAUNCH(0, float);
      else if (qm == 1) FWD_LAUNCH(1, float);
      else FWD_LAUNCH(2, float);
    }
#undef FWD_LAUNCH
#undef FWD_LAUNCH_INC
    timed_end(tev, st);
    CK(hipGetLastError());
    return 0;
  }

  // workgroups of the band per 256-column chunk and shot: kBandRows rows each
  int band_groups() const { return (band - kRim + kBandRows - 1) / kBandRows; }

  // the FP64 near-surface band of forward step n (rows [4, band) of every shot)
  BandArgs band_args(const Batch& b, const ModelArgs& m, int n, float* qa, float* qb, float* gather) const {
    BandArgs a;
    a.g = grid_args(b);
    a.m = m;
    a.band = band;
    a.dt = (double)dt;
    a.src = d_src;
    a.it = n;
    a.P = bP(n);
    a.Pm = bP(n + 1);
    a.F = bF(n);
    a.Fm = bF(n + 1);
    a.psx = bpsx(n);
    a.psxn = bpsx(n + 1);
    a.zx = bzx();
    a.Fc = fF(n);
    a.Pcf = fP(n);
    a.Po = fP(n + 1);
    a.Fo = fF(n + 1);
    a.gather = gather;
    a.qa = qa;
    a.qb = qb;
    return a;
  }

  // one forward step of a batch: one launch
  int fwd_step(const Batch& b, const ModelArgs& m, int n, int qm, float* qa, float* qb, float* gather,
               hipStream_t st) {
    return fwd_main(b, m, n, qm, qa, qb, gather, b.d_items, b.nitems, nullptr, qm == 0 ? 0 : 1, st);
  }

  // Two forward steps n, n+1 (increment form) in two launches: the core pair with step n of the FP64 band, of the
  // shell and of the core's cells next to it (they read only step n and write disjoint sets); then step n+1 of the
  // band and the shell, which read the set of step n+1 there.  Storing, the imaging terms of step n+1 lie 2 fields
  // after those of step n (qa1 = qa0 + 2 cap_cell, as every caller lays them out), and the four fields must be
  // addressable with 31-bit byte offsets (pair_store_ok).
  bool pairs(const Batch& b) const { return inc() && opt.step_pairs != 0 && !pairs_off && b.nitems_core > 0; }
  bool pair_store_ok() const { return 4.0 * 4.0 * (double)cap_cell < 2147483648.0; }
  int fwd_pair(const Batch& b, const ModelArgs& m, int n, int qm, float* qa0, float* qb0, float* qa1, float* qb1,
               float* gather, hipStream_t st) {
    if (qm != 0 && (!pair_store_ok() || qb0 != qa0 + cap_cell || qa1 != qa0 + 2 * cap_cell ||
                    qb1 != qa0 + 3 * cap_cell))
      return fail("step pair: imaging-term slots not laid out as the pair kernel addresses them");
    FwdPairArgs a;
    a.g = grid_args(b);
    a.m = m;
    a.items = b.d_items_core;
    a.nitems = b.nitems_core;
    a.C = (dt * dt) / (dh * dh);
    a.src = d_src;
    a.P0 = fP(n);
    a.F0 = fF(n);
    a.Q0 = fQ(n);
    a.R0 = fR(n);
    a.P2 = fP(n + 2);
    a.F2 = fF(n + 2);
    a.Q2 = fQ(n + 2);
    a.R2 = fR(n + 2);
    a.gather = gather;
    a.qa0 = qm != 0 ? qa0 : nullptr;
    a.qstride = (unsigned)(4 * cap_cell);
    a.it = n;
    a.sh = fwd_args(b, m, n, qa0, qb0, gather, b.d_items_shell1, b.nitems_shell1, b.d_xr_shell1, true);
    a.nshell_wg = ((b.nitems_shell1 + 3) / 4 + 7) / 8 * 8;
    dim3 grid((unsigned)(a.sh.nband_wg + a.nshell_wg + b.nitems_core));
    const int tev = timed_begin(qm == 0 ? 3 : 4, st);
    if (qm == 0)
      hipLaunchKernelGGL(k_fwd_pair<0>, grid, dim3(256), 0, st, a);
    else if (qm == 1)
      hipLaunchKernelGGL(k_fwd_pair<1>, grid, dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL(k_fwd_pair<2>, grid, dim3(256), 0, st, a);
    timed_end(tev, st);
    CK(hipGetLastError());
    return fwd_main(b, m, n + 1, qm, qa1, qb1, gather, b.d_items_shell, b.nitems_shell, b.d_xr_shell, qm == 0 ? 8 : 9,
                    st);
  }

  template <bool IMG>
  static void adj_ex_launch_img(bool grouped, int am, dim3 grid, const AdjExArgs& a, hipStream_t st) {
    if (grouped) {
      if (am == FWI_ARITH_MIXED)
        hipLaunchKernelGGL((k_adj_ex_main<double, true, IMG>), grid, dim3(256), 0, st, a);
      else if (am == FWI_ARITH_FP32D)
        hipLaunchKernelGGL((k_adj_ex_main<F32D, true, IMG>), grid, dim3(256), 0, st, a);
      else if (am == FWI_ARITH_FP32I)  // increment form
        hipLaunchKernelGGL((k_adj_ex_main<float, true, IMG, true>), grid, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((k_adj_ex_main<float, true, IMG>), grid, dim3(256), 0, st, a);
    } else {
      if (am == FWI_ARITH_MIXED)
        hipLaunchKernelGGL((k_adj_ex_main<double, false, IMG>), grid, dim3(256), 0, st, a);
      else if (am == FWI_ARITH_FP32D)
        hipLaunchKernelGGL((k_adj_ex_main<F32D, false, IMG>), grid, dim3(256), 0, st, a);
      else if (am == FWI_ARITH_FP32I)
        hipLaunchKernelGGL((k_adj_ex_main<float, false, IMG, true>), grid, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((k_adj_ex_main<float, false, IMG>), grid, dim3(256), 0, st, a);
    }
  }
  static int adj_ex_launch(bool grouped, int am, dim3 grid, const AdjExArgs& a, hipStream_t st) {
    if (a.qa)
      adj_ex_launch_img<true>(grouped, am, grid, a, st);
    else
      adj_ex_launch_img<false>(grouped, am, grid, a, st);
    return 0;
  }

  // Arguments of exact adjoint step k (m^k from m^{k+1}, m^{k+2}) over an item list, without imaging terms
  AdjExArgs adj_ex_args(const Batch& b, const ModelArgs& m, int k, const int4* items, int nitems) const {
    AdjExArgs a;
    a.g = grid_args(b);
    a.m = m;
    a.items = items;
    a.nitems = nitems;
    a.dt2 = dt * dt;
    a.Lc = adj(aset(k + 1));
    a.Lp = adj(aset(k + 2));
    a.Lo = adj(aset(k));
    a.Mc = adj(4 + aset(k + 1));
    a.Mp = adj(4 + aset(k + 2));
    a.Mo = adj(4 + aset(k));
    // C-PML memory variables of step k+1 in set (k+1) mod 2, of step k in set k mod 2
    const int cur = k & 1;
    a.sxo = adj(8 + (cur ^ 1));
    a.sxn = adj(8 + cur);
    a.zxo = adj(10 + (cur ^ 1));
    a.zxn = adj(10 + cur);
    a.szo = adj(12 + (cur ^ 1));
    a.szn = adj(12 + cur);
    a.zzo = adj(14 + (cur ^ 1));
    a.zzn = adj(14 + cur);
    a.DP = adj(16);
    a.DF = adj(17);
    a.band = opt.arith == FWI_ARITH_FP32I ? band : 0;
    a.res = resid();
    a.rit = k - 1;
    a.qa = nullptr;
    a.qb = nullptr;
    a.gh = sgrad(0);
    a.gp = sgrad(1);
    a.gn = sgrad(2);
    a.xr = nullptr;
    return a;
  }

  bool apairs_possible(const Batch& b) const {
    return opt.grad_mode == FWI_GRAD_EXACT && inc() && opt.adj_step_pairs != 0 && b.grouped;
  }
  bool apairs(const Batch& b) const { return apairs_possible(b) && !apairs_off && b.nitems_acore > 0; }
  // adjoint steps k and k-1 (mu^k, mu^{k-1} from mu^{k+1}, mu^{k+2}): the core's two steps and the shell's step k
  // in one launch, the shell's step k-1 in a second; the imaging terms of both steps are needed
  int adj_pair(const Batch& b, const ModelArgs& m, int k, const float* qa0, const float* qb0, const float* qa1,
               const float* qb1, hipStream_t st) {
    if (!qa0 || !qa1 || k < 2) return fail("adjoint step pair: needs both steps' imaging terms and k >= 2");
    AdjPairArgs a;
    a.g = grid_args(b);
    a.m = m;
    a.items = b.d_items_acore;
    a.nitems = b.nitems_acore;
    a.dt2 = dt * dt;
    a.L1 = adj(aset(k + 1));
    a.M1 = adj(4 + aset(k + 1));
    a.L2 = adj(aset(k + 2));
    a.M2 = adj(4 + aset(k + 2));
    a.L0 = adj(aset(k));
    a.M0 = adj(4 + aset(k));
    a.Lm = adj(aset(k - 1));
    a.Mm = adj(4 + aset(k - 1));
    a.res = resid();
    a.rit = k - 1;
    a.qa0 = qa0;
    a.qb0 = qb0;
    a.qa1 = qa1;
    a.qb1 = qb1;
    a.gh = sgrad(0);
    a.gp = sgrad(1);
    a.gn = sgrad(2);
    a.sh = adj_ex_args(b, m, k, b.d_items_ash, b.nitems_ash);
    a.sh.qa = qa0;
    a.sh.qb = qb0;
    a.sh.xr = b.d_xr_ash;
    a.nshell_wg = ((b.nitems_ash + 3) / 4 + 7) / 8 * 8;
    const int tev = timed_begin(5, st);
    hipLaunchKernelGGL(k_adj_pair, dim3((unsigned)(a.nshell_wg + b.nitems_acore / 4)), dim3(256), 0, st, a);
    timed_end(tev, st);
    CK(hipGetLastError());
    if (b.nitems_ash > 0) {
      AdjExArgs c = adj_ex_args(b, m, k - 1, b.d_items_ash, b.nitems_ash);
      c.qa = qa1;
      c.qb = qb1;
      c.xr = b.d_xr_ash;
      const int tes = timed_begin(7, st);
      hipLaunchKernelGGL(k_adj_shell, dim3((unsigned)(b.nitems_ash / 4)), dim3(256), 0, st, c);
      timed_end(tes, st);
      CK(hipGetLastError());
    }
    return 0;
  }

  // one adjoint step k (produces m^k from m^{k+1}, m^{k+2}); q slot pointers may be null
  int adj_step(const Batch& b, const ModelArgs& m, int k, const float* qa, const float* qb, hipStream_t st) {
    const int cur = k & 1;
    float* Lc = adj(cur);
    float* Lo = adj(cur ^ 1);
    float* Mc = adj(2 + cur);
    float* Mo = adj(2 + (cur ^ 1));
    const int am = opt.arith;
    StripArgs sa{grid_args(b), m, b.d_sblocks, b.nsblocks};
    dim3 grid((b.nitems_adj + 3) / 4);
#define ARITH_LAUNCH(KERNEL, GRID, ...)                                                     \
  do {                                                                                      \
    if (am == FWI_ARITH_MIXED)                                                              \
      hipLaunchKernelGGL(KERNEL<double>, GRID, dim3(256), 0, st, __VA_ARGS__);              \
    else if (am == FWI_ARITH_FP32D)                                                         \
      hipLaunchKernelGGL(KERNEL<F32D>, GRID, dim3(256), 0, st, __VA_ARGS__);                \
    else                                                                                    \
      hipLaunchKernelGGL(KERNEL<float>, GRID, dim3(256), 0, st, __VA_ARGS__);               \
  } while (0)
    if (opt.grad_mode == FWI_GRAD_REFCOMPAT) {
      if (b.nsblocks > 0) ARITH_LAUNCH(k_adj_ref_strip, dim3(b.nsblocks), sa, Lc, Mc, adj(4), adj(5), adj(6), adj(7));
      AdjRefArgs a;
      a.g = grid_args(b);
      a.m = m;
      a.items = b.d_items_adj;
      a.nitems = b.nitems_adj;
      a.dh2C = (dh * dh) * ((dt * dt) / (dh * dh));
      a.dt = dt;
      a.Lc = Lc;
      a.Lo = Lo;
      a.Mc = Mc;
      a.Mo = Mo;
      a.s1 = adj(4);
      a.s2 = adj(5);
      a.s3 = adj(6);
      a.s4 = adj(7);
      a.z1 = adj(8);
      a.z2 = adj(9);
      a.z3 = adj(10);
      a.z4 = adj(11);
      a.res = resid();
      a.rit = k - 1;
      a.qa = qa;
      a.qb = qb;
      a.gcx = sgrad(0);
      a.gcz = sgrad(1);
      a.gdx = sgrad(2);
      const int tev = timed_begin(2, st);
      ARITH_LAUNCH(k_adj_ref_main, grid, a);
      timed_end(tev, st);
    } else {
      AdjExArgs a = adj_ex_args(b, m, k, b.d_items_adj, b.nitems_adj);
      a.qa = qa;
      a.qb = qb;
      const int tev = timed_begin(2, st);
      if (int r = adj_ex_launch(b.grouped, am, grid, a, st)) return r;
      timed_end(tev, st);
    }
#undef ARITH_LAUNCH
    CK(hipGetLastError());
    return 0;
  }

  // the C-PML memory variables of step n (psi of its parity) to or from a checkpoint's packed strip cells
  int strip_copy(const Batch& b, int dir, int n, float* packed, hipStream_t st) {
    if (b.nsblocks == 0) return 0;
    StripArgs sa{grid_args(b), ModelArgs{nullptr, nullptr, nullptr, nx}, b.d_sblocks, b.nsblocks};
    hipLaunchKernelGGL(k_strip_copy, dim3(b.nsblocks), dim3(256), 0, st, sa, dir, fpsx(n), fpsz(n), fzx(), fzz(),
                       packed, b.nstrip);
    CK(hipGetLastError());
    return 0;
  }

  int zero_fwd(const Batch& b, hipStream_t st) {
    if (int r = zero(fwd(0), kFwdF * (size_t)cap_cell, st)) return r;
    if (band > kRim)
      if (int r = zero((float*)d_band, 2 * kBandF * (size_t)cap_band, st)) return r;
    (void)b;
    return 0;
  }
  // the band state of a checkpoint: after its cell and strip fields
  float* ckpt_band(int j) const { return (float*)(ckpt(j) + (4 * (size_t)cap_cell + 4 * (size_t)cap_strip) * ck_esize); }

  int save_ckpt(const Batch& b, int j, hipStream_t st) {
    char* c = ckpt(j);
    const int n = j * K;
    if (opt.ckpt_format != FWI_CKPT_FP32) {
      const bool h = opt.ckpt_format == FWI_CKPT_FP16;
      unsigned* am = d_ck_amax + (size_t)j * max_count * 8;
      hipLaunchKernelGGL(k_fill_u32, dim3((b.count * 8 + 255) / 256), dim3(256), 0, st, am, 0u, b.count * 8);
      hipLaunchKernelGGL(k_ck_absmax_cells, dim3(kCkBlocks, b.count), dim3(256), 0, st, b.d_shots, nz,
                         fwd(ckpt_field(0, n)), fwd(ckpt_field(1, n)), fwd(ckpt_field(2, n)), fwd(ckpt_field(3, n)), am);
      if (h)
        hipLaunchKernelGGL(k_ck_pack_cells<2>, dim3(kCkBlocks, b.count), dim3(256), 0, st,
                         b.d_shots, nz, 0, fwd(ckpt_field(0, n)), fwd(ckpt_field(1, n)), fwd(ckpt_field(2, n)),
                         fwd(ckpt_field(3, n)), (unsigned short*)c, b.ncell, am);
      else
        hipLaunchKernelGGL(k_ck_pack_cells<1>, dim3(kCkBlocks, b.count), dim3(256), 0, st,
                         b.d_shots, nz, 0, fwd(ckpt_field(0, n)), fwd(ckpt_field(1, n)), fwd(ckpt_field(2, n)),
                         fwd(ckpt_field(3, n)), (unsigned short*)c, b.ncell, am);
      if (b.nsblocks > 0) {
        StripArgs sa{grid_args(b), ModelArgs{nullptr, nullptr, nullptr, nx}, b.d_sblocks, b.nsblocks};
        hipLaunchKernelGGL(k_ck_absmax_strips, dim3(b.nsblocks), dim3(256), 0, st, sa, fpsx(n), fpsz(n), fzx(), fzz(),
                           am);
        if (h)
          hipLaunchKernelGGL(k_ck_pack_strips<2>, dim3(b.nsblocks), dim3(256), 0, st, sa, 0, fpsx(n), fpsz(n), fzx(),
                             fzz(), (unsigned short*)c + 4 * (size_t)b.ncell, b.nstrip, am);
        else
          hipLaunchKernelGGL(k_ck_pack_strips<1>, dim3(b.nsblocks), dim3(256), 0, st, sa, 0, fpsx(n), fpsz(n), fzx(),
                             fzz(), (unsigned short*)c + 4 * (size_t)b.ncell, b.nstrip, am);
      }
      CK(hipGetLastError());
      return 0;
    }
    float* cf = (float*)c;
    for (int f = 0; f < 4; ++f)
      if (int r = copy(cf + (size_t)f * b.ncell, fwd(ckpt_field(f, n)), b.ncell, st)) return r;
    if (band > kRim)
      if (int r = copy(ckpt_band(j), (const float*)d_band, 2 * kBandF * (size_t)cap_band, st)) return r;
    return strip_copy(b, 0, n, cf + 4 * (size_t)b.ncell, st);
  }
  int load_ckpt(const Batch& b, int j, hipStream_t st) {
    char* c = ckpt(j);
    const int n = j * K;
    if (opt.ckpt_format != FWI_CKPT_FP32) {
      const bool h = opt.ckpt_format == FWI_CKPT_FP16;
      const unsigned* am = d_ck_amax + (size_t)j * max_count * 8;
      if (h)
        hipLaunchKernelGGL(k_ck_pack_cells<2>, dim3(kCkBlocks, b.count), dim3(256), 0, st,
                         b.d_shots, nz, 1, fwd(ckpt_field(0, n)), fwd(ckpt_field(1, n)), fwd(ckpt_field(2, n)),
                         fwd(ckpt_field(3, n)), (unsigned short*)c, b.ncell, am);
      else
        hipLaunchKernelGGL(k_ck_pack_cells<1>, dim3(kCkBlocks, b.count), dim3(256), 0, st,
                         b.d_shots, nz, 1, fwd(ckpt_field(0, n)), fwd(ckpt_field(1, n)), fwd(ckpt_field(2, n)),
                         fwd(ckpt_field(3, n)), (unsigned short*)c, b.ncell, am);
      if (b.nsblocks > 0) {
        StripArgs sa{grid_args(b), ModelArgs{nullptr, nullptr, nullptr, nx}, b.d_sblocks, b.nsblocks};
        if (h)
          hipLaunchKernelGGL(k_ck_pack_strips<2>, dim3(b.nsblocks), dim3(256), 0, st, sa, 1, fpsx(n), fpsz(n), fzx(),
                             fzz(), (unsigned short*)c + 4 * (size_t)b.ncell, b.nstrip, am);
        else
          hipLaunchKernelGGL(k_ck_pack_strips<1>, dim3(b.nsblocks), dim3(256), 0, st, sa, 1, fpsx(n), fpsz(n), fzx(),
                             fzz(), (unsigned short*)c + 4 * (size_t)b.ncell, b.nstrip, am);
      }
      CK(hipGetLastError());
      return 0;
    }
    float* cf = (float*)c;
    for (int f = 0; f < 4; ++f)
      if (int r = copy(fwd(ckpt_field(f, n)), cf + (size_t)f * b.ncell, b.ncell, st)) return r;
    if (band > kRim)
      if (int r = copy((float*)d_band, ckpt_band(j), 2 * kBandF * (size_t)cap_band, st)) return r;
    return strip_copy(b, 1, n, cf + 4 * (size_t)b.ncell, st);
  }

  int residual(const Batch& b, const float* syn, const float* obs, float* res, double* out, int accumulate,
               hipStream_t st) {
    hipLaunchKernelGGL(k_residual, dim3(kResidualBlocks, b.count), dim3(256), 0, st, b.d_shots, nt, syn,
                       obs + b.gather_off, res, d_partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, d_partial, kResidualBlocks, b.count, 0.5, out,
                       accumulate, d_shot_misfit + b.s0);
    CK(hipGetLastError());
    return 0;
  }

  // ---------------------------------------------------------------------------------------------
  int forward_batch(const Batch& b, const ModelArgs& m, float* gather, hipStream_t st) {
    if (int r = zero_fwd(b, st)) return r;
    for (int n = 0; n < nt;) {
      if (pairs(b) && n + 1 < nt) {
        if (int r = fwd_pair(b, m, n, 0, nullptr, nullptr, nullptr, nullptr, gather, st)) return r;
        n += 2;
      } else {
        if (int r = fwd_step(b, m, n, 0, nullptr, nullptr, gather, st)) return r;
        n += 1;
      }
      if (int r = step_done()) return r;
    }
    return 0;
  }

  // obs != null: residual = syn - obs and misfit; adj_src != null: the residual is adj_src itself.
  int gradient_batch(const Batch& b, const ModelArgs& m, const float* obs, const float* adj_src, double* grad,
                     int first_batch, float* syn_out, hipStream_t st) {
    const int qm = opt.grad_mode == FWI_GRAD_EXACT ? 1 : 2;
    const int lag = opt.grad_mode == FWI_GRAD_EXACT ? 1 : 2;
    float* gather = syn_out ? syn_out + b.gather_off : gathers();
    const size_t slot = 2 * (size_t)cap_cell;
    if (int r = zero_fwd(b, st)) return r;
    // forward with checkpoints before segments 1..J-S-1; the last S segments keep their imaging terms
    const int s0 = sfirst();
    auto ckpt_at = [&](int n) { return n > 0 && n % K == 0 && n / K <= nckpt(); };
    auto qslot = [&](int n) { return n >= s0 ? qbuf() + (size_t)(n - s0) * slot : nullptr; };
    for (int n = 0; n < nt;) {
      if (ckpt_at(n))
        if (int r = save_ckpt(b, n / K, st)) return r;
      // a pair may not straddle a checkpoint or the first kept step
      if (pairs(b) && n + 1 < nt && !ckpt_at(n + 1) && (n >= s0) == (n + 1 >= s0) && (n < s0 || pair_store_ok())) {
        float* q0 = qslot(n);
        float* q1 = qslot(n + 1);
        if (int r = fwd_pair(b, m, n, n >= s0 ? qm : 0, q0, q0 ? q0 + cap_cell : nullptr, q1,
                             q1 ? q1 + cap_cell : nullptr, gather, st))
          return r;
        n += 2;
      } else {
        float* q0 = qslot(n);
        if (int r = fwd_step(b, m, n, n >= s0 ? qm : 0, q0, q0 ? q0 + cap_cell : nullptr, gather, st)) return r;
        n += 1;
      }
      if (int r = step_done()) return r;
    }
    if (adj_src) {
      if (int r = copy(resid(), adj_src + b.gather_off, b.ngather, st)) return r;
    } else {
      if (int r = residual(b, gather, obs, resid(), d_misfit, first_batch ? 0 : 1, st)) return r;
    }
    // adjoint
    if (int r = zero(adj(0), kAdjF * (size_t)cap_cell, st)) return r;
    if (int r = zero(sgrad(0), kGradF * (size_t)cap_cell, st)) return r;
    // segments below s0 are recomputed from their checkpoints into slots 0..K-1, whose kept terms the
    // adjoint has consumed by then (it walks down from nt)
    int seg = J - S;
    for (int k = nt; k >= 1;) {
      const int n = k - lag;
      const float* qa = nullptr;
      const float* qb = nullptr;
      if (n >= s0) {
        qa = qbuf() + (size_t)(n - s0) * slot;
        qb = qa + cap_cell;
      } else if (n >= 0) {
        while (n < seg * K) {
          --seg;
          if (seg > 0) {
            if (int r = load_ckpt(b, seg, st)) return r;
          } else {
            if (int r = zero_fwd(b, st)) return r;
          }
          const int e = std::min((seg + 1) * K, nt);
          for (int nn = seg * K; nn < e;) {
            float* wa = qbuf() + (size_t)(nn - seg * K) * slot;
            if (pairs(b) && nn + 1 < e && pair_store_ok()) {
              if (int r = fwd_pair(b, m, nn, qm, wa, wa + cap_cell, wa + slot, wa + slot + cap_cell, nullptr, st))
                return r;
              nn += 2;
            } else {
              if (int r = fwd_step(b, m, nn, qm, wa, wa + cap_cell, nullptr, st)) return r;
              nn += 1;
            }
            if (int r = step_done()) return r;
          }
        }
        qa = qbuf() + (size_t)(n - seg * K) * slot;
        qb = qa + cap_cell;
      }
      // a step pair where step k-1's imaging terms lie in the same kept or recomputed range
      if (qa && k >= 2 && apairs(b) && (n >= s0 ? n - 1 >= s0 : n - 1 >= seg * K)) {
        const float* qa1 = qa - slot;
        if (int r = adj_pair(b, m, k, qa, qb, qa1, qa1 + cap_cell, st)) return r;
        if (int r = step_done()) return r;
        k -= 2;
        continue;
      }
      if (int r = adj_step(b, m, k, qa, qb, st)) return r;
      if (int r = step_done()) return r;
      --k;
    }
    const long long nmodel = (long long)nx * nz;
    hipLaunchKernelGGL(k_cross_shot_sum, dim3((unsigned)((nmodel + 255) / 256)), dim3(256), 0, st, b.d_shots, b.count,
                       nx, nz, sgrad(0), sgrad(1), sgrad(2), grad, nmodel);
    CK(hipGetLastError());
    return 0;
  }

  // ---------------------------------------------------------------------------------------------
  // Pick the work-item order and rows per item of the forward kernels, then of the adjoint kernels (with
  // the imaging terms, so the gradient update is part of what is timed), by timing a few steps of each
  // candidate on the plan's own buffers and its largest batch.
  int tune(const ModelArgs& m, hipStream_t st) {
    if (tuned) return 0;
    std::vector<int> rows_c = {16, 8, 4, 24, 32, 48, 64, 96};  // 8, 4: small grids, whose few items are latency-bound
    std::vector<int> rows_a = rows_c;
    if (opt.rows_per_item > 0) rows_c = rows_a = {opt.rows_per_item};
    if (opt.adj_rows_per_item > 0) rows_a = {opt.adj_rows_per_item};
    Batch& b = batches[0];
    std::vector<int> order_f = {0, 1, 2, 3};
    if (!share_groups(b).empty()) order_f.push_back(4);
    std::vector<int> order_a = b.grouped ? std::vector<int>{0, 2, 3} : std::vector<int>{0, 1, 2, 3};
    if (opt.item_order >= 0) {
      order_f = {opt.item_order};
      order_a = {b.grouped && opt.item_order == 1 ? 0 : opt.item_order == 4 ? 3 : opt.item_order};
    }
    if (opt.adj_item_order >= 0) order_a = {b.grouped && opt.adj_item_order == 1 ? 0 : opt.adj_item_order};
    // Each candidate is timed twice over 6 steps and keeps the shorter time; the fastest candidate is taken.
    struct Cand {
      float ms;
      int order, rows;
    };
    auto pick = [&](const std::vector<Cand>& cs, int& rows, int& ord) {
      const Cand* sel = &cs[0];
      for (const Cand& c : cs)
        if (c.ms < sel->ms) sel = &c;
      rows = sel->rows;
      ord = sel->order;
    };
    // the tuner times launches with events and waits for them: not possible while the caller captures `st`
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    CK(hipStreamIsCapturing(st, &cap));
    if (cap != hipStreamCaptureStatusNone)
      return fail("the first call on a plan tunes its work items with timed launches and cannot run while the caller "
                  "captures its stream: make one call outside the capture first");
    struct Events {  // destroyed on every return path
      hipEvent_t e[2] = {nullptr, nullptr};
      ~Events() {
        for (hipEvent_t x : e)
          if (x) (void)hipEventDestroy(x);
      }
    } ev;
    CK(hipEventCreate(&ev.e[0]));
    CK(hipEventCreate(&ev.e[1]));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    const char* tlog = std::getenv("FWI_TUNE_LOG");  // "1": each candidate's time to stderr
    const bool log = tlog && tlog[0] == '1';
    auto timed = [&](auto&& body, float& ms) -> int {
      CK(hipEventRecord(e0, st));
      if (int r = body()) return r;
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&ms, e0, e1));
      return 0;
    };
    // forward
    std::vector<Cand> cands;
    for (int o : order_f)
      for (int c : rows_c) {
        if (c > nz - 8 && c != rows_c[0]) continue;
        if (int r = build_items(b, c, o)) return r;
        if (rows_c.size() == 1 && order_f.size() == 1) {
          cands.push_back({0.f, o, c});
          break;
        }
        if (int r = zero_fwd(b, st)) return r;
        for (int w = 0; w < 2; ++w)
          if (int r = fwd_step(b, m, w, 0, nullptr, nullptr, nullptr, st)) return r;
        float ms = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
          float t = 0;
          if (int r = timed([&]() -> int {
                for (int n = 2; n < 8; ++n)
                  if (int r2 = fwd_step(b, m, n, 0, nullptr, nullptr, nullptr, st)) return r2;
                return 0;
              }, t))
            return r;
          ms = std::min(ms, t);
        }
        if (log) std::fprintf(stderr, "[fwi tune] forward  order %d rows %2d: %.1f us/step\n", o, c, ms * 1e3 / 6);
        cands.push_back({ms, o, c});
      }
    pick(cands, tz, order);
    // adjoint, imaging into zeroed scratch slots: step n reads slot n mod nq, as the gradient reads a slot per step
    // (one slot for every step would let a step pair read its second step's terms from the cache)
    const int nq = std::min(qslots(), 8);
    const size_t qslot = 2 * (size_t)cap_cell;
    auto qa_of = [&](int n) { return qbuf() + (size_t)(n % nq) * qslot; };
    CK(hipMemsetAsync(qbuf(), 0, (size_t)nq * qslot * sizeof(float), st));
    cands.clear();
    for (int o : order_a)
      for (int c : rows_a) {
        if (c > nz - 8 && c != rows_a[0]) continue;
        if (int r = build_adj_items(b, c, o)) return r;
        if (rows_a.size() == 1 && order_a.size() == 1) {
          cands.push_back({0.f, o, c});
          break;
        }
        CK(hipMemsetAsync(adj(0), 0, kAdjF * (size_t)cap_cell * sizeof(float), st));
        for (int w = 0; w < 2; ++w)
          if (int r = adj_step(b, m, nt - w, qa_of(w), qa_of(w) + cap_cell, st)) return r;
        float ms = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
          float t = 0;
          if (int r = timed([&]() -> int {
                for (int n = 2; n < 8; ++n)
                  if (int r2 = adj_step(b, m, nt - n, qa_of(n), qa_of(n) + cap_cell, st)) return r2;
                return 0;
              }, t))
            return r;
          ms = std::min(ms, t);
        }
        if (log) std::fprintf(stderr, "[fwi tune] adjoint  order %d rows %2d: %.1f us/step\n", o, c, ms * 1e3 / 6);
        cands.push_back({ms, o, c});
      }
    pick(cands, tz_adj, order_adj);
    // adjoint step pairs: rows per core item, kept only where they beat single steps (six steps each)
    if (apairs_possible(b) && nt >= 10) {
      if (int r = build_adj_items(b, tz_adj, order_adj)) return r;
      std::vector<int> rows_q = {16, 32, 48, 64};
      if (opt.adj_step_pairs > 1) rows_q = {opt.adj_step_pairs};
      auto time_adj = [&](bool pr, float& ms) -> int {
        ms = 1e30f;
        CK(hipMemsetAsync(adj(0), 0, kAdjF * (size_t)cap_cell * sizeof(float), st));
        for (int w = 0; w < 2; ++w)
          if (int r = adj_step(b, m, nt - w, qa_of(w), qa_of(w) + cap_cell, st)) return r;
        for (int rep = 0; rep < 2; ++rep) {
          float t = 0;
          if (int r = timed([&]() -> int {
                for (int n = 2; n < 8; n += pr ? 2 : 1)
                  if (int r2 = pr ? adj_pair(b, m, nt - n, qa_of(n), qa_of(n) + cap_cell, qa_of(n + 1),
                                             qa_of(n + 1) + cap_cell, st)
                                  : adj_step(b, m, nt - n, qa_of(n), qa_of(n) + cap_cell, st))
                    return r2;
                return 0;
              }, t))
            return r;
          ms = std::min(ms, t);
        }
        return 0;
      };
      float best = 1e30f;
      for (int c : rows_q) {
        if (c > nz - 8 && c != rows_q[0]) continue;
        if (int r = build_adj_pair_items(b, c)) return r;
        if (rows_q.size() == 1 || b.nitems_acore == 0) {
          tz_apair = c;
          best = -1.f;
          break;
        }
        float ms = 0;
        if (int r = time_adj(true, ms)) return r;
        if (log) std::fprintf(stderr, "[fwi tune] adjoint pair rows %2d: %.1f us/step\n", c, ms * 1e3 / 6);
        if (ms < best) {
          best = ms;
          tz_apair = c;
        }
      }
      if (opt.adj_step_pairs == 1 && best >= 0.f && best < 1e29f) {
        float single = 0;
        if (int r = time_adj(false, single)) return r;
        apairs_off = best >= single;
        if (log)
          std::fprintf(stderr, "[fwi tune] adjoint pairs %s (%.1f vs %.1f us/step single)\n", apairs_off ? "off" : "on",
                       best * 1e3 / 6, single * 1e3 / 6);
      }
    }
    // fused core step pairs: rows per item
    if (inc() && opt.step_pairs != 0) {
      std::vector<int> rows_p = {16, 24, 32, 48, 64, 96};
      if (opt.step_pairs > 1) rows_p = {opt.step_pairs};
      float best = 1e30f;
      for (int c : rows_p) {
        if (c > nz - 8 && c != rows_p[0]) continue;
        if (int r = build_pair_items(b, c)) return r;
        if (rows_p.size() == 1 || b.nitems_core == 0) {
          tz_pair = c;
          break;
        }
        if (int r = zero_fwd(b, st)) return r;
        if (int r = fwd_pair(b, m, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, st)) return r;
        float ms = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
          float t = 0;
          if (int r = timed([&]() -> int {
                for (int n = 2; n < 8; n += 2)
                  if (int r2 = fwd_pair(b, m, n, 0, nullptr, nullptr, nullptr, nullptr, nullptr, st)) return r2;
                return 0;
              }, t))
            return r;
          ms = std::min(ms, t);
        }
        if (log) std::fprintf(stderr, "[fwi tune] pair     rows %2d: %.1f us/step\n", c, ms * 1e3 / 6);
        if (ms < best) {
          best = ms;
          tz_pair = c;
        }
      }
      // rows of the completing launch's shell items, with the chosen core rows
      if (rows_p.size() > 1 && b.nitems_core > 0 && best < 1e29f) {
        const int keep = tz_shell;
        int best_s = keep;
        for (int sr : {4, 8, 16}) {
          tz_shell = sr;
          if (int r = build_pair_items(b, tz_pair)) return r;
          if (int r = zero_fwd(b, st)) return r;
          if (int r = fwd_pair(b, m, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, st)) return r;
          float ms = 1e30f;
          for (int rep = 0; rep < 2; ++rep) {
            float t = 0;
            if (int r = timed([&]() -> int {
                  for (int n = 2; n < 8; n += 2)
                    if (int r2 = fwd_pair(b, m, n, 0, nullptr, nullptr, nullptr, nullptr, nullptr, st)) return r2;
                  return 0;
                }, t))
              return r;
            ms = std::min(ms, t);
          }
          if (log) std::fprintf(stderr, "[fwi tune] pair shell rows %2d: %.1f us/step\n", sr, ms * 1e3 / 6);
          if (ms < best) {
            best = ms;
            best_s = sr;
          }
        }
        tz_shell = best_s;
      }
      // tuned pairs are kept only where they beat single steps (six steps each, the chosen tilings)
      if (opt.step_pairs == 1 && best < 1e29f) {
        if (int r = build_items(b, tz, order)) return r;
        if (int r = zero_fwd(b, st)) return r;
        for (int w = 0; w < 2; ++w)
          if (int r = fwd_step(b, m, w, 0, nullptr, nullptr, nullptr, st)) return r;
        float single_ms = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
          float t = 0;
          if (int r = timed([&]() -> int {
                for (int n = 2; n < 8; ++n)
                  if (int r2 = fwd_step(b, m, n, 0, nullptr, nullptr, nullptr, st)) return r2;
                return 0;
              }, t))
            return r;
          single_ms = std::min(single_ms, t);
        }
        pairs_off = best >= single_ms;
        if (log)
          std::fprintf(stderr, "[fwi tune] pairs %s (%.1f vs %.1f us/step single)\n", pairs_off ? "off" : "on",
                       best * 1e3 / 6, single_ms * 1e3 / 6);
      }
    }
    for (auto& bb : batches) {
      if (int r = build_items(bb, tz, order)) return r;
      if (int r = build_adj_items(bb, tz_adj, order_adj)) return r;
      if (inc() && opt.step_pairs != 0)
        if (int r = build_pair_items(bb, tz_pair)) return r;
      if (apairs_possible(bb))
        if (int r = build_adj_pair_items(bb, tz_apair)) return r;
    }
    tuned = true;
    return 0;
  }
};

namespace {

int validate(const fwi_survey* s, std::string& why) {
  if (!s || !s->shots || s->nshots <= 0) return why = "survey has no shots", 1;
  if (s->nt <= 0 || s->nx < 16 || s->nz < 16) return why = "grid or time axis too small", 1;
  if (s->cpml <= 4) return why = "cpml must exceed the 4-cell stencil rim", 1;
  if (s->nz - s->cpml < 8) return why = "nz too small for the bottom C-PML", 1;
  if (!(s->dh > 0) || !(s->dt > 0)) return why = "dh and dt must be positive", 1;
  if (!s->wavelet || !s->cpml_a || !s->cpml_b) return why = "wavelet and C-PML table are required", 1;
  for (int i = 0; i < s->nshots; ++i) {
    const fwi_shot& h = s->shots[i];
    char buf[160];
    auto bad = [&](const char* w) {
      std::snprintf(buf, sizeof buf, "shot %d: %s", i, w);
      why = buf;
      return 1;
    };
    if (h.x0 < 0 || h.nx < 2 * s->cpml + 8 || h.x0 + h.nx > s->nx) return bad("window outside the model or too narrow");
    if (h.isx < 4 || h.isx > h.nx - 5 || h.isz < 4 || h.isz > s->nz - 5) return bad("source outside the interior");
    if (h.nrec < 0 || h.irz < 4 || h.irz > s->nz - 5 || (h.nrec > 0 && (h.ir0 < 4 || h.ir0 + h.nrec > h.nx - 4)))
      return bad("receivers outside the interior");
  }
  return 0;
}

}  // namespace

extern "C" {

// Gradient groups of up to gmax shots from `ord` (the batch's shots ordered by window); member lists in wave order.
constexpr int kGroupLook = 12;
static std::vector<std::vector<int>> group_shots(const std::vector<ShotDev>& sh, const std::vector<int>& ord,
                                                 int gmax) {
  // placeholder waves of a group: per 64-column band of its union window, the members whose window misses it
  auto waste = [&](const std::vector<int>& m) {
    int ux0 = 1 << 30, ux1 = 0;
    for (int s : m) {
      ux0 = std::min(ux0, sh[s].x0);
      ux1 = std::max(ux1, sh[s].x0 + sh[s].nx);
    }
    int w = 0;
    for (int c = ux0; c < ux1; c += 64)
      for (int s : m)
        if (sh[s].x0 >= c + 64 || sh[s].x0 + sh[s].nx <= c) ++w;
    return w;
  };
  std::vector<int> rest = ord;
  std::vector<std::vector<int>> out;
  while (!rest.empty()) {
    const int nc = std::min<int>(kGroupLook, (int)rest.size() - 1);
    const int k = std::min(gmax - 1, nc);
    std::vector<int> pick(k), best;
    for (int j = 0; j < k; ++j) pick[j] = j + 1;  // positions in rest, lexicographic combinations of [1, nc]
    int bw = -1;
    while (true) {
      std::vector<int> m = {rest[0]};
      for (int j : pick) m.push_back(rest[j]);
      const int w = waste(m);
      if (bw < 0 || w < bw) {
        bw = w;
        best = pick;
      }
      int j = k - 1;
      while (j >= 0 && pick[j] == nc - (k - 1 - j)) --j;
      if (j < 0) break;
      ++pick[j];
      for (int t = j + 1; t < k; ++t) pick[t] = pick[t - 1] + 1;
    }
    std::vector<int> g = {rest[0]};
    for (int j : best) g.push_back(rest[j]);
    std::vector<int> left;
    for (int j = 1; j < (int)rest.size(); ++j)
      if (std::find(best.begin(), best.end(), j) == best.end()) left.push_back(rest[j]);
    out.push_back(g);
    rest.swap(left);
  }
  return out;
}

void fwi_options_default(fwi_options* o) {
  std::memset(o, 0, sizeof *o);
  o->grad_mode = FWI_GRAD_EXACT;
  o->arith = FWI_ARITH_FP32I;
  o->item_order = -1;
  o->adj_item_order = -1;
  o->use_graph = 1;
  o->step_pairs = 1;
  o->adj_step_pairs = 1;
}

int fwi_plan_create(const fwi_survey* s, const fwi_options* opt, fwi_plan** out) {
  if (check_host("fwi_plan_create", {{out, "plan"}})) return 1;
  *out = nullptr;
  std::string why;
  if (validate(s, why)) {
    set_message(why);
    return 1;
  }
  auto* p = new fwi_plan();
  p->nx = s->nx;
  p->nz = s->nz;
  p->nt = s->nt;
  p->cpml = s->cpml;
  p->dh = s->dh;
  p->dt = s->dt;
  p->shots.assign(s->shots, s->shots + s->nshots);
  p->wavelet.assign(s->wavelet, s->wavelet + s->nt);
  p->ta.assign(s->cpml_a, s->cpml_a + s->cpml);
  p->tb.assign(s->cpml_b, s->cpml_b + s->cpml);
  if (opt)
    p->opt = *opt;
  else
    fwi_options_default(&p->opt);
  if (p->opt.grad_mode != FWI_GRAD_EXACT && p->opt.grad_mode != FWI_GRAD_REFCOMPAT) {
    delete p;
    set_message("unknown grad_mode");
    return 1;
  }
  if (p->opt.arith != FWI_ARITH_FP32 && p->opt.arith != FWI_ARITH_MIXED && p->opt.arith != FWI_ARITH_FP32D &&
      p->opt.arith != FWI_ARITH_FP32I) {
    delete p;
    set_message("unknown arith");
    return 1;
  }
  if (p->opt.ckpt_format != FWI_CKPT_FP32 && p->opt.ckpt_format != FWI_CKPT_BF16 &&
      p->opt.ckpt_format != FWI_CKPT_FP16) {
    delete p;
    set_message("unknown ckpt_format");
    return 1;
  }
  if (p->opt.ckpt_format != FWI_CKPT_FP32 && p->opt.arith != FWI_ARITH_FP32I) {
    delete p;
    set_message("16-bit checkpoints need the increment form (FWI_ARITH_FP32I): two time levels stored in 16 bits "
                "lose their difference");
    return 1;
  }
  p->ck_esize = p->opt.ckpt_format == FWI_CKPT_FP32 ? 4 : 2;
  {  // FP64 near-surface band
    const int k = p->opt.surface_rows;
    const bool allowed = p->opt.arith == FWI_ARITH_FP32I && p->opt.ckpt_format == FWI_CKPT_FP32;
    if (k == 0) {
      p->band = allowed && s->nz - s->cpml - 4 >= 24 ? 24 : 0;
    } else if (k > 0) {
      if (!allowed || k <= kRim || k > s->nz - s->cpml - 4) {
        delete p;
        set_message("surface_rows needs FWI_ARITH_FP32I, FP32 checkpoints and 4 < surface_rows <= nz - cpml - 4");
        return 1;
      }
      p->band = k;
    }
  }
  if (p->opt.grad_group < 0 || p->opt.grad_group > 4) {
    delete p;
    set_message("grad_group must be 0..4");
    return 1;
  }
  if (p->opt.item_order < -1 || p->opt.item_order > 4 || p->opt.adj_item_order < -1 || p->opt.adj_item_order > 3 ||
      p->opt.rows_per_item < 0 || p->opt.adj_rows_per_item < 0 || p->opt.step_pairs < 0 || p->opt.adj_step_pairs < 0) {
    delete p;
    set_message("item_order must be -1..4, adj_item_order -1..3, rows_per_item, adj_rows_per_item, step_pairs and "
                "adj_step_pairs >= 0");
    return 1;
  }
  const int nt = p->nt;
  p->K = p->opt.ckpt_interval > 0 ? std::min(p->opt.ckpt_interval, nt) : (int)std::ceil(std::sqrt((double)nt));
  p->J = (nt + p->K - 1) / p->K;
  p->tz = p->tz_adj = p->opt.rows_per_item > 0 ? p->opt.rows_per_item : 32;
  if (p->opt.adj_rows_per_item > 0) p->tz_adj = p->opt.adj_rows_per_item;

  // per-shot sizes
  const int ns = s->nshots;
  std::vector<long long> cells(ns), strips(ns), gath(ns);
  for (int i = 0; i < ns; ++i) {
    cells[i] = (long long)row_pitch(p->shots[i].nx) * p->nz;
    strips[i] = strip_count(p->shots[i].nx, p->nz, p->cpml);
    gath[i] = (long long)nt * p->shots[i].nrec;
  }
  // per-cell floats: the fixed forward, adjoint and gradient fields + 2K imaging terms + 4 per checkpoint (2 in 16 bits)
  const double ckf = (double)p->ck_esize / 4.0;  // checkpoint values per float
  const double per_cell = (double)fwi_plan::kFixedF + 2.0 * p->K + 4.0 * ckf * (p->J - 1);
  const double per_strip = 4.0 * ckf * (p->J - 1);
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
    delete p;
    set_message("hipMemGetInfo failed (no GPU?)");
    return 1;
  }
  double budget = p->opt.mem_budget > 0 ? p->opt.mem_budget : 0.85 * (double)freeb;
  // greedy batches in shot order
  int i = 0;
  long long gather_off = 0;
  while (i < ns) {
    Batch b;
    b.s0 = i;
    double bytes = 0;
    while (i < ns) {
      double add = 4.0 * (per_cell * cells[i] + per_strip * strips[i] + 2.0 * gath[i]);
      if (b.count > 0 && (bytes + add > budget || (p->opt.max_batch > 0 && b.count >= p->opt.max_batch))) break;
      bytes += add;
      b.ncell += cells[i];
      b.nstrip += strips[i];
      b.ngather += gath[i];
      ++b.count;
      ++i;
    }
    b.gather_off = gather_off;
    gather_off += b.ngather;
    p->batches.push_back(std::move(b));
  }
  p->gather_total = gather_off;
  for (auto& b : p->batches) {
    p->cap_cell = std::max(p->cap_cell, b.ncell);
    p->cap_strip = std::max(p->cap_strip, b.nstrip);
    p->cap_gather = std::max(p->cap_gather, b.ngather);
  }
  p->cap_band = (long long)p->band * (p->cap_cell / p->nz);  // band rows x the row pitches of the largest batch
  // trailing segments kept from the forward: as many as the budget allows with FP32 checkpoints (recomputation
  // reproduces them exactly, so the result does not change), one with 16-bit checkpoints (whose recomputed
  // segments start from rounded states), or exactly store_segments
  if (p->opt.store_segments > 0) {
    p->S = std::min(p->opt.store_segments, p->J);
  } else if (p->opt.ckpt_format == FWI_CKPT_FP32) {
    p->S = 1;
    while (p->S < p->J && 4.0 * (double)p->workspace_floats(p->S + 1) <= budget) ++p->S;
  } else {
    p->S = 1;
  }
  const size_t nfloat = p->workspace_floats(p->S);
  p->work_bytes = nfloat * sizeof(float);
  int rc = 0;
  auto ck = [&](hipError_t e, const char* w) {
    if (e != hipSuccess && rc == 0) rc = p->fail(e, w);
  };
  ck(hipMalloc(&p->d_work, p->work_bytes), "hipMalloc(workspace)");
  if (p->cap_band > 0)
    ck(hipMalloc(&p->d_band, fwi_plan::kBandF * (size_t)p->cap_band * sizeof(double)), "hipMalloc(band)");
  ck(hipMalloc(&p->d_src, nt * sizeof(float)), "hipMalloc(src)");
  ck(hipMalloc(&p->d_ta, p->cpml * sizeof(float)), "hipMalloc(ta)");
  ck(hipMalloc(&p->d_tb, p->cpml * sizeof(float)), "hipMalloc(tb)");
  for (auto& b : p->batches) p->max_count = std::max(p->max_count, b.count);
  const int max_count = p->max_count;
  ck(hipMalloc(&p->d_partial, (size_t)kResidualBlocks * max_count * sizeof(double)), "hipMalloc(partial)");
  if (p->opt.ckpt_format != FWI_CKPT_FP32)
    ck(hipMalloc(&p->d_ck_amax, (size_t)p->J * max_count * 8 * sizeof(unsigned)), "hipMalloc(ckpt scales)");
  ck(hipMalloc(&p->d_misfit, sizeof(double)), "hipMalloc(misfit)");
  ck(hipMalloc(&p->d_shot_misfit, (size_t)ns * sizeof(double)), "hipMalloc(shot misfit)");
  if (rc == 0) ck(hipMemcpy(p->d_src, p->wavelet.data(), nt * sizeof(float), hipMemcpyHostToDevice), "copy src");
  if (rc == 0) ck(hipMemcpy(p->d_ta, p->ta.data(), p->cpml * sizeof(float), hipMemcpyHostToDevice), "copy ta");
  if (rc == 0) ck(hipMemcpy(p->d_tb, p->tb.data(), p->cpml * sizeof(float), hipMemcpyHostToDevice), "copy tb");
  if (rc == 0) ck(hipMemset(p->d_work, 0, p->work_bytes), "zero workspace");
  // per-batch tables
  for (auto& b : p->batches) {
    if (rc) break;
    long long off = 0, soff = 0, goff = 0, boff = 0;
    for (int k = 0; k < b.count; ++k) {
      const fwi_shot& h = p->shots[b.s0 + k];
      ShotDev d;
      d.nx = h.nx;
      d.ps = row_pitch(h.nx);
      d.x0 = h.x0;
      d.isx = h.isx;
      d.isz = h.isz;
      d.irz = h.irz;
      d.ir0 = h.ir0;
      d.nrec = h.nrec;
      d.nstrip = strip_count(h.nx, p->nz, p->cpml);
      d.off = off;
      d.soff = soff;
      d.goff = goff;
      d.gslot = off;
      d.gx0 = h.x0;
      d.gnx = h.nx;
      d.gps = d.ps;
      d.lead = 1;
      d.boff = boff;
      boff += (long long)d.ps * p->band;
      off += (long long)d.ps * p->nz;
      b.nreal += (long long)h.nx * p->nz;
      soff += d.nstrip;
      goff += (long long)nt * h.nrec;
      b.h_shots.push_back(d);
      for (int k0 = 0; k0 < d.nstrip; k0 += 256) b.h_sblocks.push_back(make_int2(k, k0));
    }
    b.nsblocks = (int)b.h_sblocks.size();
    // gradient groups (exact mode): the batch's shots ordered by window (x0, width, then batch order); each group
    // is led by the first shot not yet grouped and completed by the up to grad_group-1 shots, among the next
    // kGroupLook ungrouped ones in that order, whose union window leaves the fewest placeholder waves (64-column
    // bands of the union a member's window does not reach; first such choice in lexicographic order, so shots
    // with identical windows group in runs).  Each group's FP32 slot spans the union of its windows, slots back
    // to back (their total is at most the batch's cells).
    const int gmax = p->opt.grad_mode == FWI_GRAD_EXACT ? (p->opt.grad_group == 0 ? 4 : p->opt.grad_group) : 1;
    std::vector<int> ord(b.count);
    for (int k = 0; k < b.count; ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](int u, int v) {
      const ShotDev &a = b.h_shots[u], &c = b.h_shots[v];
      return a.x0 != c.x0 ? a.x0 < c.x0 : a.nx < c.nx;
    });
    std::vector<std::vector<int>> gsets = group_shots(b.h_shots, ord, gmax);
    long long slot = 0;
    for (const std::vector<int>& gm : gsets) {
      Batch::Group g;
      g.n = (int)gm.size();
      int ux1 = 0;
      g.ux0 = 1 << 30;
      for (int j = 0; j < g.n; ++j) {
        g.m[j] = gm[j];
        const ShotDev& d = b.h_shots[g.m[j]];
        g.ux0 = std::min(g.ux0, d.x0);
        ux1 = std::max(ux1, d.x0 + d.nx);
      }
      g.unx = ux1 - g.ux0;
      for (int j = 0; j < g.n; ++j) {
        ShotDev& d = b.h_shots[g.m[j]];
        d.gslot = slot;
        d.gx0 = g.ux0;
        d.gnx = g.unx;
        d.gps = row_pitch(g.unx);
        d.lead = j == 0;
      }
      slot += (long long)row_pitch(g.unx) * p->nz;
      if (g.n > 1) b.grouped = true;
      b.groups.push_back(g);
    }
    ck(hipMalloc(&b.d_shots, b.h_shots.size() * sizeof(ShotDev)), "hipMalloc(shots)");
    if (rc == 0)
      ck(hipMemcpy(b.d_shots, b.h_shots.data(), b.h_shots.size() * sizeof(ShotDev), hipMemcpyHostToDevice), "copy shots");
    if (rc == 0 && b.nsblocks > 0) {
      ck(hipMalloc(&b.d_sblocks, b.h_sblocks.size() * sizeof(int2)), "hipMalloc(sblocks)");
      if (rc == 0)
        ck(hipMemcpy(b.d_sblocks, b.h_sblocks.data(), b.h_sblocks.size() * sizeof(int2), hipMemcpyHostToDevice),
           "copy sblocks");
    }
    // until the first call tunes (or applies the pinned tiling), the items are a placeholder tiling
    p->order = p->opt.item_order > 0 ? p->opt.item_order : 0;
    const int ao = p->opt.adj_item_order >= 0 ? p->opt.adj_item_order
                   : p->opt.item_order >= 2    ? std::min(p->opt.item_order, 3)
                                               : 0;
    p->order_adj = b.grouped && ao == 1 ? 0 : ao;
    if (rc == 0) rc = p->build_items(b, p->tz, p->order);
    if (rc == 0) rc = p->build_adj_items(b, p->tz_adj, p->order_adj);
    if (rc == 0 && p->inc() && p->opt.step_pairs != 0) rc = p->build_pair_items(b, p->tz_pair);
    if (rc == 0 && p->apairs_possible(b)) rc = p->build_adj_pair_items(b, p->tz_apair);
  }
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return 0;
}

void fwi_plan_destroy(fwi_plan* p) { delete p; }

int fwi_plan_set_wavelet(fwi_plan* p, const float* w) {
  if (check_host("fwi_plan_set_wavelet", {{p, "plan"}, {w, "wavelet"}})) return 1;
  p->wavelet.assign(w, w + p->nt);
  // launches and graph replays of this plan may still run on the caller's (possibly non-blocking) streams and
  // read the table: wait for the device before overwriting it
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(p->d_src, w, p->nt * sizeof(float), hipMemcpyHostToDevice);
  return e == hipSuccess ? 0 : p->fail(e, "set_wavelet");
}

int fwi_plan_set_cpml(fwi_plan* p, const float* a, const float* b) {
  if (check_host("fwi_plan_set_cpml", {{p, "plan"}, {a, "cpml_a"}, {b, "cpml_b"}})) return 1;
  p->ta.assign(a, a + p->cpml);
  p->tb.assign(b, b + p->cpml);
  hipError_t e = hipDeviceSynchronize();  // as in fwi_plan_set_wavelet
  if (e == hipSuccess) e = hipMemcpy(p->d_ta, a, p->cpml * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_tb, b, p->cpml * sizeof(float), hipMemcpyHostToDevice);
  return e == hipSuccess ? 0 : p->fail(e, "set_cpml");
}

long long fwi_plan_gather_size(const fwi_plan* p) { return p->gather_total; }
int fwi_plan_batch_size(const fwi_plan* p) { return p->batches.empty() ? 0 : p->batches[0].count; }
int fwi_plan_batch_count(const fwi_plan* p) { return (int)p->batches.size(); }
double fwi_plan_workspace_bytes(const fwi_plan* p) { return (double)p->work_bytes; }

int fwi_forward(fwi_plan* p, const float* vh, const float* vp, const float* vn, float* gathers, void* stream) {
  if (check_host("fwi_forward", {{p, "plan"}}) ||
      check_arrays("fwi_forward", {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}, {gathers, "gathers"}}))
    return 1;
  hipStream_t st = (hipStream_t)stream;
  ModelArgs m{vh, vp, vn, p->nx};
  if (int r = p->tune(m, st)) return r;
  for (size_t i = 0; i < p->batches.size(); ++i) {
    Batch& b = p->batches[i];
    const std::array<const void*, 10> key{(const void*)1, (const void*)i, vh, vp, vn, gathers};
    if (int r = p->run_graph(key, st, [&](hipStream_t s) { return p->forward_batch(b, m, gathers + b.gather_off, s); }))
      return r;
  }
  p->timing_collect();
  return 0;
}

int fwi_misfit(fwi_plan* p, const float* vh, const float* vp, const float* vn, const float* obs, double* misfit,
               void* stream) {
  if (check_host("fwi_misfit", {{p, "plan"}, {misfit, "misfit"}}) ||
      check_arrays("fwi_misfit", {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}, {obs, "obs"}}))
    return 1;
  hipStream_t st = (hipStream_t)stream;
  ModelArgs m{vh, vp, vn, p->nx};
  if (int r = p->tune(m, st)) return r;
  for (size_t i = 0; i < p->batches.size(); ++i) {
    Batch& b = p->batches[i];
    const int acc = i > 0;
    const std::array<const void*, 10> key{(const void*)2, (const void*)i, vh, vp, vn, obs};
    if (int r = p->run_graph(key, st, [&](hipStream_t s) {
          if (int r2 = p->forward_batch(b, m, p->gathers(), s)) return r2;
          return p->residual(b, p->gathers(), obs, nullptr, p->d_misfit, acc, s);
        }))
      return r;
  }
  hipError_t e = hipMemcpyAsync(misfit, p->d_misfit, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? 0 : p->fail(e, "misfit readback");
}

int fwi_gradient(fwi_plan* p, const float* vh, const float* vp, const float* vn, const float* obs, double* grad,
                 double* misfit, float* syn, void* stream) {
  if (check_host("fwi_gradient", {{p, "plan"}}) ||
      check_arrays("fwi_gradient", {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}, {obs, "obs"}, {grad, "grad"}}) ||
      (syn && check_arrays("fwi_gradient", {{syn, "syn"}})))
    return 1;
  hipStream_t st = (hipStream_t)stream;
  ModelArgs m{vh, vp, vn, p->nx};
  if (int r = p->tune(m, st)) return r;
  hipError_t e = hipMemsetAsync(grad, 0, 3 * (size_t)p->nx * p->nz * sizeof(double), st);
  if (e != hipSuccess) return p->fail(e, "zero grad");
  for (size_t i = 0; i < p->batches.size(); ++i) {
    Batch& b = p->batches[i];
    const int first = i == 0;
    const std::array<const void*, 10> key{(const void*)3, (const void*)i, vh, vp, vn, obs, grad, syn};
    if (int r = p->run_graph(key, st, [&](hipStream_t s) {
          return p->gradient_batch(b, m, obs, nullptr, grad, first, syn, s);
        }))
      return r;
  }
  if (misfit) {
    e = hipMemcpyAsync(misfit, p->d_misfit, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return p->fail(e, "misfit readback");
  }
  p->timing_collect();
  return 0;
}

int fwi_plan_shot_misfits(fwi_plan* p, double* out, void* stream) {
  if (check_host("fwi_plan_shot_misfits", {{p, "plan"}, {out, "out"}})) return 1;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(out, p->d_shot_misfit, p->shots.size() * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? 0 : p->fail(e, "shot misfit readback");
}

void fwi_plan_set_timing(fwi_plan* p, int stride) {
  p->timing_collect();
  p->timing_stride = stride;
  for (int k = 0; k < fwi_plan::kKinds; ++k) {
    p->kt_ms[k] = 0;
    p->kt_count[k] = 0;
    p->launch_no[k] = 0;
  }
}

long long fwi_plan_kernel_launches(const fwi_plan* p, int kind) {
  if (kind < 0 || kind >= fwi_plan::kKinds) return -1;
  return p->launch_no[kind];
}

long long fwi_plan_core_cells(const fwi_plan* p) {
  if (p->batches.empty() || !p->pairs(p->batches[0])) return 0;
  long long c = 0;
  const int za = p->pair_za(), zb = p->pair_zb();
  for (auto& b : p->batches)
    for (const ShotDev& d : b.h_shots) {
      const int x0 = pair_cx0(p->cpml), x1 = pair_cx1(d.nx, p->cpml);
      if (x1 > x0 && zb > za) c += (long long)(x1 - x0) * (zb - za);
    }
  return c;
}

int fwi_plan_kernel_time(fwi_plan* p, int kind, double* total_ms, long long* count) {
  if (kind < 0 || kind >= fwi_plan::kKinds) return 1;
  p->timing_collect();
  *total_ms = p->kt_ms[kind];
  *count = p->kt_count[kind];
  return 0;
}

long long fwi_plan_cells(const fwi_plan* p, long long* strip_cells) {
  long long c = 0, s = 0;
  for (auto& b : p->batches) {
    c += b.nreal;
    s += b.nstrip;
  }
  if (strip_cells) *strip_cells = s;
  return c;
}

int fwi_plan_ckpt_interval(const fwi_plan* p) { return p->K; }
int fwi_plan_kept_steps(const fwi_plan* p) { return p->nt - p->sfirst(); }
int fwi_plan_rows_per_item(const fwi_plan* p) { return p->tz; }
int fwi_plan_item_order(const fwi_plan* p) { return p->order; }
int fwi_plan_graph_stats(const fwi_plan* p, long long* captures, long long* replays) {
  if (captures) *captures = p->graph_captures;
  if (replays) *replays = p->graph_replays;
  return 0;
}
int fwi_plan_graph_nodes(const fwi_plan* p, long long* largest, long long* last) {
  if (largest) *largest = p->graph_nodes_largest;
  if (last) *last = p->graph_nodes_last;
  return 0;
}
long long fwi_plan_adj_core_cells(const fwi_plan* p) {
  long long c = 0;
  if (!p->batches.empty() && p->apairs(p->batches[0]))
    for (auto& b : p->batches) c += b.acore_cells;
  return c;
}
int fwi_plan_adj_pair_rows(const fwi_plan* p) {
  return (!p->batches.empty() && p->apairs(p->batches[0])) ? p->tz_apair : 0;
}
int fwi_plan_pair_rows(const fwi_plan* p) {
  return (!p->batches.empty() && p->pairs(p->batches[0])) ? p->tz_pair : 0;
}
int fwi_plan_surface_rows(const fwi_plan* p) { return p->band > kRim ? p->band : 0; }
int fwi_plan_adj_tiling(const fwi_plan* p, int* rows, int* order, int* grouped) {
  if (rows) *rows = p->tz_adj;
  if (order) *order = p->order_adj;
  if (grouped) *grouped = p->batches.empty() ? 0 : (int)p->batches[0].grouped;
  return 0;
}

int fwi_gradient_src(fwi_plan* p, const float* vh, const float* vp, const float* vn, const float* adj_src,
                     double* grad, float* syn, void* stream) {
  if (check_host("fwi_gradient_src", {{p, "plan"}}) ||
      check_arrays("fwi_gradient_src",
                   {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}, {adj_src, "adj_src"}, {grad, "grad"}}) ||
      (syn && check_arrays("fwi_gradient_src", {{syn, "syn"}})))
    return 1;
  hipStream_t st = (hipStream_t)stream;
  ModelArgs m{vh, vp, vn, p->nx};
  if (int r = p->tune(m, st)) return r;
  hipError_t e = hipMemsetAsync(grad, 0, 3 * (size_t)p->nx * p->nz * sizeof(double), st);
  if (e != hipSuccess) return p->fail(e, "zero grad");
  for (size_t i = 0; i < p->batches.size(); ++i) {
    Batch& b = p->batches[i];
    const std::array<const void*, 10> key{(const void*)4, (const void*)i, vh, vp, vn, adj_src, grad, syn};
    if (int r = p->run_graph(key, st, [&](hipStream_t s) {
          return p->gradient_batch(b, m, nullptr, adj_src, grad, 0, syn, s);
        }))
      return r;
  }
  p->timing_collect();
  return 0;
}

int fwi_chain_rule(const float* vp, const float* eps, const float* dl, const double* g, double* out, long long n,
                   void* stream) {
  if (n <= 0) return 0;
  if (check_arrays("fwi_chain_rule", {{vp, "vp"}, {eps, "eps"}, {dl, "delta"}, {g, "grad"}, {out, "out"}})) return 1;
  hipLaunchKernelGGL(k_chain_rule, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, vp, eps, dl,
                     g, out, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_message(std::string("chain_rule: ") + hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

}  // extern "C"
