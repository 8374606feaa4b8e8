// Host side of the persistent LD sweep (kernel and design notes in sweep_persistent.h): the per-XCD task queues of a
// list of view grids, the workspace layout, the launch and the completion check.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "citrus_hip.h"
#include "common.h"
#include "sweep_persistent.h"

namespace citrus {
namespace {

// Clears the tickets, the status word and the done flags of the next sweep (a kernel, not a memset, so that the
// sweep can be captured in a graph).
__global__ void persistent_clear_kernel(int* ctrl, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctrl[i] = 0;
}


hipError_t launch_persistent(int nsh, bool bf16, const PersistentArgs& a, int layout, size_t lds, hipStream_t s) {
  if (nsh == 16) return persistent_launch_p3(a, bf16, layout, lds, s);
  if (nsh == 36) return persistent_launch_p5(a, bf16, layout, lds, s);
  return hipErrorInvalidValue;
}

// Grid fields and workspace offsets of every view (pointers unset); false with a message on a bad shape.
bool build_views(int nviews, const int* shapes, int G, std::vector<PersistentView>& views, int64_t& flag_ints,
                 int64_t& face_floats_per_dc, std::string& why) {
  constexpr int B = 4;
  int64_t flag_off = 0, face_off = 0;
  for (int v = 0; v < nviews; ++v) {
    const int R = shapes[3 * v], C = shapes[3 * v + 1], S = shapes[3 * v + 2];
    if (R < 1 || C < 1 || S < 1 || (int64_t)R * C * S > (1 << 26)) {
      why = "view grid shapes must be positive and hold at most 2^26 cells";
      return false;
    }
    PersistentView pv{};
    pv.R = R; pv.C = C; pv.S = S; pv.N = R * C * S;
    pv.nbr = (R + B - 1) / B; pv.nbc = (C + B - 1) / B; pv.nbs = (S + B - 1) / B;
    pv.K = (pv.nbr - 1) + (pv.nbc - 1) + (pv.nbs - 1);
    const int64_t nb = (int64_t)pv.nbr * pv.nbc * pv.nbs;
    pv.flag_off = flag_off;
    pv.face_off = face_off;  // in units of dc until the run scales it
    flag_off += (int64_t)G * 8 * nb;
    face_off += 2 * (int64_t)G * nb * 3 * 48;
    views.push_back(pv);
  }
  flag_ints = flag_off;
  face_floats_per_dc = face_off;
  return true;
}

// The task queues: (view, group) domains go to XCD queues longest first onto the least loaded queue; each queue
// lists its tasks by (pass, block plane, octant of the pass, domain, block) -- a topological order of the
// dependencies of for_each_dependency, whose tasks all lie in the same domain and so in the same queue.
void build_queues(const std::vector<PersistentView>& views, int G, std::vector<int4>& table, std::vector<int>& qoff) {
  const int nviews = (int)views.size();
  std::vector<std::pair<int64_t, int>> doms;
  for (int v = 0; v < nviews; ++v)
    for (int g = 0; g < G; ++g) doms.emplace_back((int64_t)views[v].nbr * views[v].nbc * views[v].nbs, v * 256 + g);
  std::stable_sort(doms.begin(), doms.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
  std::vector<std::vector<int>> qdoms(kQueues);
  std::vector<int64_t> load(kQueues, 0);
  for (const auto& d : doms) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    qdoms[q].push_back(d.second);
    load[q] += d.first;
  }
  // blocks of each (view, octant) by block plane, in the launch sequence's block order
  std::vector<std::vector<std::vector<std::vector<int4>>>> byplane(nviews);
  int Kmax = 0;
  for (int v = 0; v < nviews; ++v) {
    const auto& s = views[v];
    Kmax = std::max(Kmax, s.K);
    byplane[v].assign(8, std::vector<std::vector<int4>>(s.K + 1));
    for (int o = 0; o < 8; ++o)
      for (int bi = 0; bi < s.nbr; ++bi)
        for (int bj = 0; bj < s.nbc; ++bj)
          for (int bk = 0; bk < s.nbs; ++bk)
            byplane[v][o][block_plane_of(o, bi, bj, bk, s.nbr, s.nbc, s.nbs)].push_back(make_int4(0, bi, bj, bk));
  }
  table.clear();
  qoff.assign(1, 0);
  for (int q = 0; q < kQueues; ++q) {
    std::vector<int> ds = qdoms[q];
    std::sort(ds.begin(), ds.end());
    for (int pass = 0; pass < 4; ++pass)
      for (int d = 0; d <= Kmax; ++d)
        for (int which = 0; which < 2; ++which) {
          const int o = which == 0 ? pass : 7 - pass;
          for (int dom : ds) {
            const int v = dom / 256, g = dom % 256;
            if (d > views[v].K) continue;
            for (const int4& b : byplane[v][o][d]) table.push_back(make_int4(v | (g << 8) | (o << 16), b.y, b.z, b.w));
          }
        }
    qoff.push_back((int)table.size());
  }
}

std::string check_plan_args(int nviews, const int* shapes, int G) {
  if (nviews < 1 || nviews > kMaxViews) return "nviews must be in 1.." + std::to_string(kMaxViews);
  if (!shapes) return "shapes is NULL";
  if (G < 1 || G > 255) return "G must be in 1..255";
  return "";
}

}  // namespace
}  // namespace citrus

// ======================================================================================================================
// plan
// ======================================================================================================================

struct CitrusPersistentPlan {
  int nviews = 0, G = 0;
  std::vector<citrus::PersistentView> shapes;  // grid fields and workspace offsets (pointers unset)
  std::vector<int> qoff;                       // kQueues + 1
  int64_t flag_ints = 0;
  int64_t face_floats_per_dc = 0;              // face workspace per direction of the chunk
  int64_t ntasks = 0;
  int4* d_tasks = nullptr;
  int* d_qoff = nullptr;
};

extern "C" int citrus_persistent_plan_create(int nviews, const int* shapes, int G, CitrusPersistentPlan** out) {
  using namespace citrus;
  if (!out) return fail(CITRUS_ERR_ARG, "plan pointer is NULL");
  *out = nullptr;
  std::string why = check_plan_args(nviews, shapes, G);
  if (!why.empty()) return fail(CITRUS_ERR_ARG, why);
  auto* plan = new CitrusPersistentPlan();
  plan->nviews = nviews;
  plan->G = G;
  if (!build_views(nviews, shapes, G, plan->shapes, plan->flag_ints, plan->face_floats_per_dc, why)) {
    delete plan;
    return fail(CITRUS_ERR_ARG, why);
  }
  std::vector<int4> table;
  build_queues(plan->shapes, G, table, plan->qoff);
  plan->ntasks = (int64_t)table.size();
  if (plan->ntasks >= (int64_t)1 << 31) {
    delete plan;
    return fail(CITRUS_ERR_ARG, "too many sweep tasks for one persistent plan");
  }
  hipError_t e = hipMalloc(&plan->d_tasks, sizeof(int4) * table.size());
  if (e == hipSuccess) e = hipMalloc(&plan->d_qoff, sizeof(int) * plan->qoff.size());
  if (e == hipSuccess) e = hipMemcpy(plan->d_tasks, table.data(), sizeof(int4) * table.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(plan->d_qoff, plan->qoff.data(), sizeof(int) * plan->qoff.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (plan->d_tasks) (void)hipFree(plan->d_tasks);
    if (plan->d_qoff) (void)hipFree(plan->d_qoff);
    delete plan;
    return fail(CITRUS_ERR_HIP, std::string("persistent plan upload: ") + hipGetErrorString(e));
  }
  *out = plan;
  return CITRUS_OK;
}

extern "C" void citrus_persistent_plan_destroy(CitrusPersistentPlan* plan) {
  if (!plan) return;
  if (plan->d_tasks) (void)hipFree(plan->d_tasks);
  if (plan->d_qoff) (void)hipFree(plan->d_qoff);
  delete plan;
}

extern "C" int64_t citrus_persistent_schedule_table(int nviews, const int* shapes, int G, int* tasks, int* qoff,
                                                    int* deps) {
  using namespace citrus;
  std::string why = check_plan_args(nviews, shapes, G);
  if (!why.empty()) return fail(CITRUS_ERR_ARG, why), -1;
  std::vector<PersistentView> views;
  int64_t flag_ints = 0, face = 0;
  if (!build_views(nviews, shapes, G, views, flag_ints, face, why)) return fail(CITRUS_ERR_ARG, why), -1;
  std::vector<int4> table;
  std::vector<int> q;
  build_queues(views, G, table, q);
  const int64_t n = (int64_t)table.size();
  if (qoff)
    for (int i = 0; i <= kQueues; ++i) qoff[i] = q[i];
  if (tasks)
    for (int64_t i = 0; i < n; ++i) {
      tasks[4 * i] = table[i].x; tasks[4 * i + 1] = table[i].y; tasks[4 * i + 2] = table[i].z; tasks[4 * i + 3] = table[i].w;
    }
  if (deps) {
    // index of task (view, group, octant, block) in the table
    std::vector<int64_t> where(flag_ints, -1);
    auto slot = [&](int v, int g, int o, int i, int j, int k) {
      const PersistentView& pv = views[v];
      const int64_t nb = (int64_t)pv.nbr * pv.nbc * pv.nbs;
      return pv.flag_off + ((int64_t)g * 8 + o) * nb + ((int64_t)i * pv.nbc + j) * pv.nbs + k;
    };
    for (int64_t i = 0; i < n; ++i) {
      const int4 e = table[i];
      where[slot(e.x & 0xff, (e.x >> 8) & 0xff, (e.x >> 16) & 7, e.y, e.z, e.w)] = i;
    }
    for (int64_t i = 0; i < n; ++i) {
      const int4 e = table[i];
      const int v = e.x & 0xff, g = (e.x >> 8) & 0xff, o = (e.x >> 16) & 7;
      const PersistentView& pv = views[v];
      int m = 0;
      for_each_dependency(pv.nbr, pv.nbc, pv.nbs, pv.K, o, e.y, e.z, e.w, [&](int oo, int a, int b, int c) {
        deps[7 * i + m++] = (int)where[slot(v, g, oo, a, b, c)];
      });
      for (; m < 7; ++m) deps[7 * i + m] = -1;
    }
  }
  return n;
}

namespace {
int64_t flags_bytes(const CitrusPersistentPlan* plan) {
  return ((int64_t)(citrus::kCtrlInts + plan->flag_ints) * 4 + 255) / 256 * 256;
}
}  // namespace

extern "C" int64_t citrus_persistent_workspace_bytes(const CitrusPersistentPlan* plan, int dc) {
  if (!plan || dc < 1 || dc > 32) return -1;
  return flags_bytes(plan) + plan->face_floats_per_dc * dc * 4;
}

extern "C" int64_t citrus_persistent_num_tasks(const CitrusPersistentPlan* plan) { return plan ? plan->ntasks : -1; }

extern "C" int citrus_persistent_sweep_run(const CitrusPersistentPlan* plan, const CitrusPersistentArgs* args,
                                           void* workspace, int64_t workspace_bytes, void* stream) {
  using namespace citrus;
  if (!plan || !args || !workspace) return fail(CITRUS_ERR_ARG, "plan, args or workspace is NULL");
  const CitrusPersistentArgs& A = *args;
  if (A.G != plan->G) return fail(CITRUS_ERR_ARG, "G differs from the plan's");
  if (A.dc < 1 || A.dc > 32) return fail(CITRUS_ERR_ARG, "dc must be in 1..32");
  if (A.d0 < 0 || A.d0 + A.dc > A.D8) return fail(CITRUS_ERR_ARG, "chunk [d0, d0+dc) outside the octant");
  if (A.nsh != 16 && A.nsh != 36)
    return fail(CITRUS_ERR_UNSUPPORTED, "the persistent sweep is built for nsh 16 and 36 (P3, P5)");
  if (!A.Y || !A.Yw || !A.views) return fail(CITRUS_ERR_ARG, "NULL direction table or view array");
  if (A.source_layout != -1 && (A.source_layout < 1 || A.source_layout > 3))
    return fail(CITRUS_ERR_ARG, "source_layout must be -1 (no source), 1, 2 or 3");
  if (A.source_layout == 3 && !A.fc_omega) return fail(CITRUS_ERR_ARG, "source_layout 3 needs fc_omega");
  if (workspace_bytes < citrus_persistent_workspace_bytes(plan, A.dc))
    return fail(CITRUS_ERR_ARG, "workspace smaller than citrus_persistent_workspace_bytes(plan, dc)");
  if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0) return fail(CITRUS_ERR_ARG, "workspace must be 256-B aligned");
  const bool bf16 = A.views[0].T_bf16 != nullptr;
  if (bf16 && A.nsh != 16 && A.nsh != 36)
    return fail(CITRUS_ERR_UNSUPPORTED, "the bfloat16 contraction is built for P3 and P5 (nsh 16 and 36)");
  PersistentArgs pa{};
  pa.G = A.G; pa.D8 = A.D8; pa.d0 = A.d0; pa.dc = A.dc; pa.M = 8 * A.D8;
  if (A.spin_limit < 0) return fail(CITRUS_ERR_ARG, "spin_limit must be >= 0 (0: the default)");
  pa.nviews = plan->nviews; pa.accumulate = A.accumulate ? 1 : 0;
  pa.spin_limit = A.spin_limit > 0 ? A.spin_limit : kSpinLimit;
  if (A.waves_per_cu < 0) return fail(CITRUS_ERR_ARG, "waves_per_cu must be >= 0");
  pa.waves_per_cu = A.waves_per_cu;
  pa.sink = A.status;
  pa.Y = A.Y; pa.Yw = A.Yw;
  pa.fc_omega = A.source_layout == 3 ? A.fc_omega : nullptr;
  pa.tasks = plan->d_tasks; pa.qoff = plan->d_qoff;
  pa.ctrl = static_cast<int*>(workspace);
  pa.flags = pa.ctrl + kCtrlInts;
  pa.faces = reinterpret_cast<float*>(static_cast<char*>(workspace) + flags_bytes(plan));
  const int words = A.source_layout == 1 ? 4 : 1;
  for (int v = 0; v < plan->nviews; ++v) {
    const CitrusSweepView& w = A.views[v];
    PersistentView pv = plan->shapes[v];
    if (!w.mu_t || !w.phi || !w.cdir) return fail(CITRUS_ERR_ARG, "view " + std::to_string(v) + ": NULL buffer");
    if ((w.T_bf16 != nullptr) != bf16 || (bf16 ? (!w.T_low || w.T) : !w.T))
      return fail(CITRUS_ERR_ARG, "every view needs float32 T, or every view the mixed operand (T_bf16 and T_low)");
    if ((A.source_layout < 0) != (w.source == nullptr))
      return fail(CITRUS_ERR_ARG, "view " + std::to_string(v) + ": source does not match source_layout");
    if (A.source_layout == 3 && (!w.fc_scale || !w.fc_beta || !w.fc_u))
      return fail(CITRUS_ERR_ARG, "view " + std::to_string(v) + ": source_layout 3 needs fc_scale, fc_beta and fc_u");
    if ((int64_t)pv.N * A.nsh * 16 >= (int64_t)kDrop || (int64_t)pv.N * pa.M * 4 * words >= (int64_t)kDrop)
      return fail(CITRUS_ERR_UNSUPPORTED, "view " + std::to_string(v) + ": grid too large for 32-bit buffer offsets");
    pv.face_off *= A.dc;
    pv.mu = w.mu_t;
    pv.T = bf16 ? w.T_bf16 : static_cast<const void*>(w.T);
    pv.Tlo = w.T_low;
    pv.src = w.source;
    pv.phi = w.phi;
    pv.cdir = w.cdir;
    pv.fc_scale = A.source_layout == 3 ? w.fc_scale : nullptr;
    pv.fc_beta = A.source_layout == 3 ? w.fc_beta : nullptr;
    pv.fc_u = A.source_layout == 3 ? w.fc_u : nullptr;
    pa.views[v] = pv;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int64_t nclear = kCtrlInts + plan->flag_ints;
  hipLaunchKernelGGL(persistent_clear_kernel, dim3((unsigned)((nclear + 255) / 256)), dim3(256), 0, s, pa.ctrl,
                     nclear);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = launch_persistent(A.nsh, bf16, pa, A.source_layout, mfma_lds_bytes(A.dc), s);
  if (e == hipErrorInvalidValue) return fail(CITRUS_ERR_UNSUPPORTED, "the persistent sweep is built for nsh 16 and 36");
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("persistent sweep launch: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

extern "C" int citrus_persistent_check(const CitrusPersistentPlan* plan, const void* workspace, int* status,
                                       int64_t* unfinished, void* stream) {
  using namespace citrus;
  if (!plan || !workspace || !status || !unfinished) return fail(CITRUS_ERR_ARG, "NULL argument");
  std::vector<int> host(kCtrlInts + plan->flag_ints);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(host.data(), workspace, host.size() * sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("persistent check: ") + hipGetErrorString(e));
  *status = host[kStatusWord];
  int64_t missing = 0;
  for (int64_t i = 0; i < plan->flag_ints; ++i) missing += host[kCtrlInts + i] == 0;
  *unfinished = missing;
  return CITRUS_OK;
}
