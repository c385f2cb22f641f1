// torch_pfc.cpp -- libtorch autograd::Function over the plain PFC trajectory.
//
//   psi_T = pfc_trajectory(psi0, eps, nsteps, dt, spacing, every, mixed, snaps)
//
// forward runs nsteps fused ETD1 steps (pfc_hip_pfc_step) from psi0 with the
// operators of eps and keeps checkpoints of the field; backward receives
// dJ/dpsi_T, recomputes the states it needs from the checkpoints and applies
// the adjoint step (pfc_hip_pfc_adjoint_step) from the last step back to the
// first, accumulating dJ/deps.  It returns dJ/dpsi0 and dJ/deps.
//
// Two checkpoint schedules:
//   snaps == 0  uniform: a checkpoint every `every` steps; the backward
//               recomputes one segment at a time and holds its states.
//               Memory: ceil(nsteps/every) checkpoints + `every` states.
//   snaps >= 1  binomial (the optimal schedule of reversal with a fixed
//               number of snapshots, Griewank's "revolve"): besides the
//               initial field at most `snaps` fields are held at any time,
//               and the number of recomputed steps is the minimum for that
//               budget.  The split points come from the exact recurrence
//                 T(1, s) = 0,  T(n, 0) = n(n-1)/2,
//                 T(n, s) = min_{0<m<n} m + T(n-m, s-1) + T(m, s),
//               (T = forward steps of the reversal of n steps with the
//               state at their start held and s free slots), solved by
//               dynamic programming; the forward pass stores the
//               checkpoints along the schedule's first chain, so the
//               backward recomputes T(nsteps, snaps) minus that chain.
// Both schedules apply the same adjoint steps to the same states in the same
// order, so their gradients agree bit for bit.
//
// psi0 is a contiguous float64 device tensor of shape (nz, ny, nx); eps is a
// float64 scalar tensor.  mixed = true runs the FP32-mixed mode: the field is
// stepped as its FP32 deviation d = psi - m from the exact mean m (kept in
// double, conserved), and the adjoint runs in FP32 at psi = m + d; psi_T and
// the gradients are returned in float64.  Every library call is enqueued on
// torch's current HIP stream, the stream torch's own kernels of this
// function (clones, casts, means) run on.
#include <torch/extension.h>

#include <c10/hip/HIPStream.h>

#include "../../include/pfc_hip.h"

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace {

void check(int rc, const char *what) {
  if (rc != PFC_HIP_OK)
    throw std::runtime_error(std::string(what) + ": " + pfc_hip_last_error());
}

void *cur_stream() { return (void *)c10::hip::getCurrentHIPStream().stream(); }

struct Plan {
  pfc_hip_plan_t h = nullptr;
  ~Plan() {
    if (h) pfc_hip_plan_destroy(h);
  }
};

// one plan per grid shape and precision, kept for the process
pfc_hip_plan_t plan_for(int nx, int ny, int nz, int prec) {
  static std::map<std::tuple<int, int, int, int>, std::unique_ptr<Plan>> plans;
  auto &p = plans[{nx, ny, nz, prec}];
  if (!p) {
    p = std::make_unique<Plan>();
    check(pfc_hip_plan_create(nx, ny, nz, prec, 0, &p->h), "pfc_trajectory: plan");
  }
  return p->h;
}

// mixed: the mean-split step around `mean`; otherwise the plain step
void set_ops(pfc_hip_plan_t plan, const std::vector<double> &spacing, double eps, double dt,
             bool mixed, double mean) {
  if (spacing.size() != 3) throw std::invalid_argument("pfc_trajectory: spacing needs 3 values");
  const double params[4] = {spacing[0], spacing[1], spacing[2], eps};
  check(pfc_hip_pfc_set_operators(plan, PFC_HIP_MODEL_PFC, params, 4, dt),
        "pfc_trajectory: operators");
  check(pfc_hip_pfc_set_mean(plan, mixed ? 1 : 0, mixed ? mean : 0.0),
        "pfc_trajectory: step mode");
}

void step(pfc_hip_plan_t plan, torch::Tensor &psi, int64_t n) {
  if (n > 0)
    check(pfc_hip_pfc_step(plan, psi.data_ptr(), (int)n, cur_stream()), "pfc_trajectory: step");
}

void check_field(const torch::Tensor &psi, const char *what) {
  if (!psi.is_cuda() || psi.scalar_type() != torch::kFloat64 || psi.dim() != 3 ||
      !psi.is_contiguous())
    throw std::invalid_argument(std::string(what) +
                                ": psi must be a contiguous float64 device tensor (nz, ny, nx)");
}

// ---------------------------------------------------------------------------
// Binomial checkpoint schedule.
// ---------------------------------------------------------------------------
class Schedule {
public:
  Schedule(int64_t n, int64_t s) : n_(n), s_(std::min<int64_t>(s, n)) {
    if (n < 1 || s < 0) throw std::invalid_argument("checkpoint schedule: nsteps >= 1, snaps >= 0");
    if ((double)n * (double)n * (double)(s_ + 1) > 2e9)
      throw std::invalid_argument("checkpoint schedule: nsteps^2 * (snaps + 1) > 2e9");
    cost_.assign((size_t)(s_ + 1) * (size_t)(n_ + 1), -1);
    split_.assign(cost_.size(), 0);
    for (int64_t m = 1; m <= n_; ++m)
      for (int64_t k = 0; k <= s_; ++k) solve(m, k);
  }
  // forward steps of the reversal of m steps with k free slots
  int64_t cost(int64_t m, int64_t k) const { return cost_[idx(m, std::min(k, s_))]; }
  // first checkpoint position (steps from the start) of that reversal, 0 if none
  int64_t split(int64_t m, int64_t k) const { return split_[idx(m, std::min(k, s_))]; }

private:
  size_t idx(int64_t m, int64_t k) const { return (size_t)k * (size_t)(n_ + 1) + (size_t)m; }
  void solve(int64_t m, int64_t k) {
    int64_t best, at = 0;
    if (m == 1) {
      best = 0;
    } else if (k == 0) {
      best = m * (m - 1) / 2;
    } else {
      best = INT64_MAX;
      for (int64_t j = 1; j < m; ++j) {
        const int64_t c = j + cost_[idx(m - j, k - 1)] + cost_[idx(j, k)];
        if (c < best) { // first minimum: the smallest split
          best = c;
          at = j;
        }
      }
    }
    cost_[idx(m, k)] = best;
    split_[idx(m, k)] = at;
  }
  int64_t n_, s_;
  std::vector<int64_t> cost_, split_;
};

// Positions the forward pass checkpoints: the chain of first splits from 0.
std::vector<int64_t> forward_chain(const Schedule &sc, int64_t n, int64_t s) {
  std::vector<int64_t> pos{0};
  int64_t at = 0, left = n, k = s;
  while (left > 1 && k > 0) {
    const int64_t j = sc.split(left, k);
    if (j <= 0) break;
    at += j;
    left -= j;
    --k;
    pos.push_back(at);
  }
  return pos;
}

struct Reverser {
  pfc_hip_plan_t plan;
  const Schedule &sc;
  torch::Tensor &lam;
  double &dr;
  std::map<int64_t, torch::Tensor> &ck; // position -> state held
  int64_t recomputed = 0;

  // Reverse steps [start, start + m): ck holds the state at `start`; k more
  // states may be held while this runs.
  void run(int64_t start, int64_t m, int64_t k) {
    if (m == 1) {
      check(pfc_hip_pfc_adjoint_step(plan, ck.at(start).data_ptr(), lam.data_ptr(), &dr,
                                     cur_stream()),
            "pfc_trajectory: adjoint step");
      return;
    }
    const int64_t j = k > 0 ? sc.split(m, k) : 0;
    if (j == 0) { // no slot: recompute from the start for every step, last first
      for (int64_t i = m - 1; i >= 0; --i) {
        torch::Tensor psi = ck.at(start).clone();
        step(plan, psi, i);
        recomputed += i;
        check(pfc_hip_pfc_adjoint_step(plan, psi.data_ptr(), lam.data_ptr(), &dr, cur_stream()),
              "pfc_trajectory: adjoint step");
      }
      return;
    }
    const int64_t mid = start + j;
    const bool had = ck.count(mid) != 0; // stored by the forward pass
    if (!had) {
      torch::Tensor psi = ck.at(start).clone();
      step(plan, psi, j);
      recomputed += j;
      ck[mid] = psi;
    }
    run(mid, m - j, k - 1);
    ck.erase(mid);
    run(start, j, k);
  }
};

class PfcTrajectory : public torch::autograd::Function<PfcTrajectory> {
public:
  static torch::Tensor forward(torch::autograd::AutogradContext *ctx, torch::Tensor psi0,
                               torch::Tensor eps, int64_t nsteps, double dt,
                               std::vector<double> spacing, int64_t every, bool mixed,
                               int64_t snaps) {
    check_field(psi0, "pfc_trajectory");
    if (nsteps < 1 || every < 1 || snaps < 0)
      throw std::invalid_argument("pfc_trajectory: nsteps, every >= 1 and snaps >= 0");
    const int nz = (int)psi0.size(0), ny = (int)psi0.size(1), nx = (int)psi0.size(2);
    pfc_hip_plan_t plan = plan_for(nx, ny, nz, mixed ? PFC_HIP_F32 : PFC_HIP_F64);
    const double e = eps.item<double>();
    const double mean = mixed ? psi0.mean().item<double>() : 0.0;
    set_ops(plan, spacing, e, dt, mixed, mean);
    std::vector<torch::Tensor> ckpts;
    std::vector<int64_t> where;
    if (snaps > 0) {
      const Schedule sc(nsteps, snaps);
      where = forward_chain(sc, nsteps, snaps);
    } else {
      for (int64_t s = 0; s < nsteps; s += every) where.push_back(s);
    }
    torch::Tensor psi = mixed ? (psi0 - mean).to(torch::kFloat32).contiguous() : psi0.clone();
    int64_t at = 0;
    for (int64_t p : where) {
      step(plan, psi, p - at);
      at = p;
      ckpts.push_back(psi.clone());
    }
    step(plan, psi, nsteps - at);
    ctx->saved_data["ckpts"] = ckpts;
    ctx->saved_data["where"] = where;
    ctx->saved_data["nsteps"] = nsteps;
    ctx->saved_data["every"] = every;
    ctx->saved_data["snaps"] = snaps;
    ctx->saved_data["dt"] = dt;
    ctx->saved_data["eps"] = e;
    ctx->saved_data["spacing"] = spacing;
    ctx->saved_data["eps_options"] = eps;
    ctx->saved_data["mixed"] = mixed;
    ctx->saved_data["mean"] = mean;
    return mixed ? psi.to(torch::kFloat64) + mean : psi;
  }

  static torch::autograd::tensor_list backward(torch::autograd::AutogradContext *ctx,
                                               torch::autograd::tensor_list grads) {
    auto ckpts = ctx->saved_data["ckpts"].toTensorVector();
    const auto where = ctx->saved_data["where"].toIntVector();
    const int64_t nsteps = ctx->saved_data["nsteps"].toInt();
    const int64_t every = ctx->saved_data["every"].toInt();
    const int64_t snaps = ctx->saved_data["snaps"].toInt();
    const double dt = ctx->saved_data["dt"].toDouble();
    const double e = ctx->saved_data["eps"].toDouble();
    const auto spacing = ctx->saved_data["spacing"].toDoubleVector();
    const torch::Tensor eps_like = ctx->saved_data["eps_options"].toTensor();
    const bool mixed = ctx->saved_data["mixed"].toBool();
    const double mean = ctx->saved_data["mean"].toDouble();
    torch::Tensor lam = grads[0].to(torch::kFloat64).contiguous().clone();
    check_field(lam, "pfc_trajectory backward");
    if (mixed) lam = lam.to(torch::kFloat32).contiguous();
    const int nz = (int)lam.size(0), ny = (int)lam.size(1), nx = (int)lam.size(2);
    pfc_hip_plan_t plan = plan_for(nx, ny, nz, mixed ? PFC_HIP_F32 : PFC_HIP_F64);
    set_ops(plan, spacing, e, dt, mixed, mean);
    double dr = 0.0;
    if (snaps > 0) {
      const Schedule sc(nsteps, snaps);
      std::map<int64_t, torch::Tensor> ck;
      for (size_t i = 0; i < where.size(); ++i) ck[where[i]] = ckpts[i];
      ckpts.clear();
      Reverser rv{plan, sc, lam, dr, ck};
      rv.run(0, nsteps, snaps);
    } else {
      for (int64_t seg = (int64_t)ckpts.size() - 1; seg >= 0; --seg) {
        const int64_t s0 = seg * every, s1 = std::min(s0 + every, nsteps);
        std::vector<torch::Tensor> states;
        torch::Tensor psi = ckpts[seg].clone();
        states.push_back(psi.clone());
        for (int64_t s = s0 + 1; s < s1; ++s) {
          step(plan, psi, 1);
          states.push_back(psi.clone());
        }
        for (int64_t s = s1 - 1; s >= s0; --s)
          check(pfc_hip_pfc_adjoint_step(plan, states[s - s0].data_ptr(), lam.data_ptr(), &dr,
                                         cur_stream()),
                "pfc_trajectory: adjoint step");
      }
    }
    torch::Tensor deps = torch::full({}, dr, eps_like.options());
    return {lam.to(torch::kFloat64), deps, torch::Tensor(), torch::Tensor(), torch::Tensor(),
            torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

torch::Tensor pfc_trajectory(torch::Tensor psi0, torch::Tensor eps, int64_t nsteps, double dt,
                             std::vector<double> spacing, int64_t every, bool mixed,
                             int64_t snaps) {
  return PfcTrajectory::apply(psi0, eps, nsteps, dt, spacing, every, mixed, snaps);
}

// (recomputed steps of the backward, peak fields held besides the initial
// one) of a schedule: snaps == 0 the uniform one with `every`, else binomial.
std::tuple<int64_t, int64_t> schedule_cost(int64_t nsteps, int64_t every, int64_t snaps) {
  if (nsteps < 1 || every < 1 || snaps < 0)
    throw std::invalid_argument("schedule_cost: nsteps, every >= 1 and snaps >= 0");
  if (snaps == 0) {
    const int64_t segs = (nsteps + every - 1) / every;
    const int64_t last = nsteps - (segs - 1) * every;
    const int64_t full = segs - 1;
    // a segment of length L recomputes L-1 steps and holds its L states
    return {full * (every - 1) + (last - 1), (segs - 1) + std::min(every, nsteps)};
  }
  const Schedule sc(nsteps, snaps);
  const auto chain = forward_chain(sc, nsteps, snaps);
  // the chain's advances are done by the forward pass
  return {sc.cost(nsteps, snaps) - chain.back(), std::min(snaps, nsteps)};
}

// T(m, snaps) for m = 1..nsteps (the recurrence above)
std::vector<int64_t> binomial_costs(int64_t nsteps, int64_t snaps) {
  const Schedule sc(nsteps, snaps);
  std::vector<int64_t> out;
  for (int64_t m = 1; m <= nsteps; ++m) out.push_back(sc.cost(m, snaps));
  return out;
}

} // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "libtorch autograd wrapper of the MI355X PFC trajectory (forward and adjoint)";
  m.def("pfc_trajectory", &pfc_trajectory, "psi_T of nsteps PFC ETD1 steps (differentiable)",
        pybind11::arg("psi0"), pybind11::arg("eps"), pybind11::arg("nsteps"),
        pybind11::arg("dt"), pybind11::arg("spacing"), pybind11::arg("every") = 20,
        pybind11::arg("mixed") = false, pybind11::arg("snaps") = 0);
  m.def("schedule_cost", &schedule_cost,
        "(recomputed backward steps, fields held) of the uniform (snaps=0) or binomial schedule",
        pybind11::arg("nsteps"), pybind11::arg("every") = 20, pybind11::arg("snaps") = 0);
  m.def("binomial_costs", &binomial_costs,
        "forward steps of the optimal reversal of m = 1..nsteps steps with the start held "
        "and snaps slots",
        pybind11::arg("nsteps"), pybind11::arg("snaps"));
}
