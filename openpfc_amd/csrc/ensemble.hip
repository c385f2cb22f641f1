// ensemble.hip -- independent PFC runs of uneven sizes advanced together.
//
// Runs that share extents, spacing and dt form a group; each group is one
// batched plan (pfc_hip_plan_create_batched) whose fields sit back to back
// in the caller's buffer, with per-run eps and mean in device memory.  One
// step of the whole ensemble (every group's five passes, the groups on
// parallel branches) is captured once into a hipGraph; each step is then a
// single hipGraphLaunch.  The step never allocates and issues no memset, so
// the captured graph replays the eager step exactly.
//
// No reference counterpart: the reference advances one Model per Simulator
// (include/openpfc/frontend/simulator.hpp:549-553), one step at a time
// (docs/api/examples/10_complete_pfc_model.cpp:121-144).
#include "../../include/pfc_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

namespace pfc_amd {
extern thread_local std::string g_last_error; // pfc_hip_last_error()
unsigned plan_generation(const pfc_hip_plan_s *p); // pfc_hip.hip
}

namespace {

constexpr int kParams = 6;    // dx, dy, dz, eps, dt, mean per run
constexpr int kMaxBranches = 8; // parallel branches of the captured step

int fail(int code, const std::string &msg) {
  pfc_amd::g_last_error = msg;
  return code;
}

#define ENS_TRY(expr, what)                                                    \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(PFC_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Group {
  int ext[3];
  double sp[3], dt;
  std::vector<int> runs; // in run order
  size_t offset = 0;     // first value of the group's fields
  pfc_hip_plan_t plan = nullptr;
};

bool same_key(const Group &g, const int *e, const double *p) {
  return g.ext[0] == e[0] && g.ext[1] == e[1] && g.ext[2] == e[2] && g.sp[0] == p[0] &&
         g.sp[1] == p[1] && g.sp[2] == p[2] && g.dt == p[4];
}

// Groups in order of their first run, with the offsets of every run.
int build_layout(int nruns, const int *ext, const double *params, std::vector<Group> &groups,
                 std::vector<size_t> &offsets, size_t &total) {
  if (nruns < 1 || !ext || !params)
    return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble: need nruns >= 1, extents and params");
  groups.clear();
  offsets.assign((size_t)nruns, 0);
  for (int r = 0; r < nruns; ++r) {
    const int *e = ext + 3 * r;
    const double *p = params + kParams * r;
    if (e[0] < 2 || e[1] < 1 || e[2] < 1)
      return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble: bad extents of run " + std::to_string(r));
    for (int i = 0; i < kParams; ++i)
      if (!std::isfinite(p[i]))
        return fail(PFC_HIP_EINVAL,
                    "pfc_hip_ensemble: parameter " + std::to_string(i) + " of run " +
                        std::to_string(r) + " not finite");
    if (!(p[0] > 0) || !(p[1] > 0) || !(p[2] > 0))
      return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble: spacing of run " + std::to_string(r));
    size_t g = 0;
    while (g < groups.size() && !same_key(groups[g], e, p)) ++g;
    if (g == groups.size()) {
      Group ng{};
      for (int i = 0; i < 3; ++i) {
        ng.ext[i] = e[i];
        ng.sp[i] = p[i];
      }
      ng.dt = p[4];
      groups.push_back(ng);
    }
    groups[g].runs.push_back(r);
  }
  total = 0;
  for (Group &g : groups) {
    g.offset = total;
    const size_t per = (size_t)g.ext[0] * g.ext[1] * g.ext[2];
    for (size_t m = 0; m < g.runs.size(); ++m) offsets[(size_t)g.runs[m]] = total + m * per;
    total += per * g.runs.size();
  }
  return 0;
}

} // namespace

struct pfc_hip_ensemble_s {
  int nruns = 0, prec = 0, mean_split = 0, device = 0;
  size_t esize = 8;
  char *fields = nullptr;
  std::vector<Group> groups;
  std::vector<size_t> offsets;
  hipStream_t cap = nullptr;
  std::vector<hipStream_t> branch;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned> cap_gen; // group plans' generations at the capture
};

namespace {

void drop_graph(pfc_hip_ensemble_s *e) {
  if (e->exec) {
    (void)hipDeviceSynchronize(); // replays still in flight finish first
    (void)hipGraphExecDestroy(e->exec);
  }
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

void release(pfc_hip_ensemble_s *e) {
  drop_graph(e);
  for (Group &g : e->groups)
    if (g.plan) pfc_hip_plan_destroy(g.plan);
  for (hipEvent_t v : e->join)
    if (v) (void)hipEventDestroy(v);
  for (hipStream_t s : e->branch)
    if (s) (void)hipStreamDestroy(s);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->cap) (void)hipStreamDestroy(e->cap);
}

int set_device(pfc_hip_ensemble_s *e) {
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur == e->device) return 0;
  ENS_TRY(hipSetDevice(e->device), "hipSetDevice");
  return 0;
}

// Record one step of every group on the branch streams, between a fork on
// `s` and the joins back onto it.
int enqueue_step(pfc_hip_ensemble_s *e, hipStream_t s) {
  ENS_TRY(hipEventRecord(e->fork, s), "ensemble fork");
  const size_t nb = e->branch.size();
  for (size_t b = 0; b < nb; ++b) ENS_TRY(hipStreamWaitEvent(e->branch[b], e->fork, 0), "fork wait");
  for (size_t g = 0; g < e->groups.size(); ++g) {
    const Group &gr = e->groups[g];
    if (int rc = pfc_hip_pfc_step(gr.plan, e->fields + gr.offset * e->esize, 1, e->branch[g % nb]))
      return rc;
  }
  for (size_t b = 0; b < nb; ++b) {
    ENS_TRY(hipEventRecord(e->join[b], e->branch[b]), "ensemble join");
    ENS_TRY(hipStreamWaitEvent(s, e->join[b], 0), "join wait");
  }
  return 0;
}

int capture(pfc_hip_ensemble_s *e) {
  drop_graph(e);
  ENS_TRY(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal), "begin capture");
  int rc = enqueue_step(e, e->cap);
  hipGraph_t g = nullptr;
  hipError_t he = hipStreamEndCapture(e->cap, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (he != hipSuccess) return fail(PFC_HIP_EDEVICE, std::string("end capture: ") + hipGetErrorString(he));
  e->graph = g;
  he = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
  if (he != hipSuccess) {
    e->exec = nullptr;
    drop_graph(e);
    return fail(PFC_HIP_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(he));
  }
  e->cap_gen.clear();
  for (const Group &gr : e->groups) e->cap_gen.push_back(pfc_amd::plan_generation(gr.plan));
  return 0;
}

// Whether a group plan changed a launch argument of its step since the
// capture (an eps or mean held as a scalar, the mean-split mode, a variant):
// the graph would replay the old values.
bool graph_stale(const pfc_hip_ensemble_s *e) {
  if (e->cap_gen.size() != e->groups.size()) return true;
  for (size_t g = 0; g < e->groups.size(); ++g)
    if (pfc_amd::plan_generation(e->groups[g].plan) != e->cap_gen[g]) return true;
  return false;
}

int check_ens(pfc_hip_ensemble_t e) {
  if (!e) return fail(PFC_HIP_EINVAL, "null ensemble");
  return 0;
}

} // namespace

extern "C" {

int pfc_hip_ensemble_layout(int nruns, const int *ext, const double *params, size_t *offsets,
                            size_t *total) {
  std::vector<Group> groups;
  std::vector<size_t> off;
  size_t t = 0;
  if (int rc = build_layout(nruns, ext, params, groups, off, t)) return rc;
  if (offsets)
    for (int r = 0; r < nruns; ++r) offsets[r] = off[(size_t)r];
  if (total) *total = t;
  return 0;
}

int pfc_hip_ensemble_create(int nruns, const int *ext, const double *params, int precision,
                            int mean_split, void *fields, pfc_hip_ensemble_t *ens) {
  if (!ens) return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble_create: ens is null");
  *ens = nullptr;
  if (!fields) return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble_create: null fields");
  if (precision != PFC_HIP_F64 && precision != PFC_HIP_F32)
    return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble_create: unknown precision");
  auto *e = new (std::nothrow) pfc_hip_ensemble_s();
  if (!e) return fail(PFC_HIP_ENOMEM, "pfc_hip_ensemble_create: host allocation");
  size_t total = 0;
  int rc = build_layout(nruns, ext, params, e->groups, e->offsets, total);
  e->nruns = nruns;
  e->prec = precision;
  e->mean_split = mean_split ? 1 : 0;
  e->esize = precision == PFC_HIP_F64 ? 8 : 4;
  e->fields = (char *)fields;
  if (!rc && hipGetDevice(&e->device) != hipSuccess)
    rc = fail(PFC_HIP_EDEVICE, "pfc_hip_ensemble_create: no HIP device");
  for (size_t g = 0; g < e->groups.size() && !rc; ++g) {
    Group &gr = e->groups[g];
    const int nb = (int)gr.runs.size();
    rc = pfc_hip_plan_create_batched(gr.ext[0], gr.ext[1], gr.ext[2], nb, precision, 0,
                                     &gr.plan);
    std::vector<double> eps((size_t)nb), mean((size_t)nb);
    for (int m = 0; m < nb; ++m) {
      const double *p = params + kParams * gr.runs[(size_t)m];
      eps[(size_t)m] = p[3];
      mean[(size_t)m] = p[5];
    }
    const double op[4] = {gr.sp[0], gr.sp[1], gr.sp[2], eps[0]};
    if (!rc) rc = pfc_hip_pfc_set_operators(gr.plan, PFC_HIP_MODEL_PFC, op, 4, gr.dt);
    if (!rc) rc = pfc_hip_pfc_set_members(gr.plan, eps.data(), mean.data(), e->mean_split);
  }
  // default-priority streams: the capture origin and one branch per group
  // (at most kMaxBranches; further groups share them round-robin)
  const int nbr = (int)std::min<size_t>(e->groups.size(), kMaxBranches);
  if (!rc && hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking) != hipSuccess)
    rc = fail(PFC_HIP_EDEVICE, "pfc_hip_ensemble_create: stream");
  if (!rc && hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
    rc = fail(PFC_HIP_EDEVICE, "pfc_hip_ensemble_create: event");
  for (int b = 0; b < nbr && !rc; ++b) {
    hipStream_t s = nullptr;
    hipEvent_t v = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&v, hipEventDisableTiming) != hipSuccess)
      rc = fail(PFC_HIP_EDEVICE, "pfc_hip_ensemble_create: branch stream");
    if (s) e->branch.push_back(s);
    if (v) e->join.push_back(v);
  }
  if (rc) {
    release(e);
    delete e;
    return rc;
  }
  *ens = e;
  return 0;
}

int pfc_hip_ensemble_destroy(pfc_hip_ensemble_t ens) {
  if (!ens) return 0;
  (void)set_device(ens);
  (void)hipDeviceSynchronize();
  release(ens);
  delete ens;
  return 0;
}

int pfc_hip_ensemble_groups(pfc_hip_ensemble_t ens, int *ngroups) {
  if (int rc = check_ens(ens)) return rc;
  if (!ngroups) return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble_groups: null");
  *ngroups = (int)ens->groups.size();
  return 0;
}

int pfc_hip_ensemble_group(pfc_hip_ensemble_t ens, int group, pfc_hip_plan_t *plan, int *batch,
                           size_t *offset) {
  if (int rc = check_ens(ens)) return rc;
  if (group < 0 || group >= (int)ens->groups.size())
    return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble_group: group out of range");
  const Group &g = ens->groups[(size_t)group];
  if (plan) *plan = g.plan;
  if (batch) *batch = (int)g.runs.size();
  if (offset) *offset = g.offset;
  return 0;
}

int pfc_hip_ensemble_autotune(pfc_hip_ensemble_t ens, int reps, void *stream) {
  if (int rc = check_ens(ens)) return rc;
  if (reps < 1) return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble_autotune: reps < 1");
  if (int rc = set_device(ens)) return rc;
  for (Group &g : ens->groups)
    if (int rc = pfc_hip_pfc_autotune(g.plan, ens->fields + g.offset * ens->esize, reps, stream,
                                      nullptr, 0))
      return rc;
  drop_graph(ens); // variants changed: capture again at the next step
  return 0;
}

int pfc_hip_ensemble_step(pfc_hip_ensemble_t ens, int nsteps, int use_graph, void *stream) {
  if (int rc = check_ens(ens)) return rc;
  if (nsteps < 0) return fail(PFC_HIP_EINVAL, "pfc_hip_ensemble_step: nsteps < 0");
  if (int rc = set_device(ens)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (!use_graph) {
    for (int i = 0; i < nsteps; ++i)
      if (int rc = enqueue_step(ens, s)) return rc;
    return 0;
  }
  if (nsteps > 0 && (!ens->exec || graph_stale(ens)))
    if (int rc = capture(ens)) return rc;
  for (int i = 0; i < nsteps; ++i) ENS_TRY(hipGraphLaunch(ens->exec, s), "hipGraphLaunch");
  return 0;
}

} // extern "C"
