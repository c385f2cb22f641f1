// torch_pfc.cpp -- libtorch autograd::Function over the plain PFC trajectory.
//
//   psi_T = pfc_trajectory(psi0, eps, nsteps, dt, spacing, every, mixed)
//
// forward runs nsteps fused ETD1 steps (pfc_hip_pfc_step) from psi0 with
// the operators of eps, keeping a checkpoint of the field every `every`
// steps.  backward receives dJ/dpsi_T and walks the segments from the last
// one: it recomputes the states of the segment from its checkpoint and
// applies the adjoint step (pfc_hip_pfc_adjoint_step) from each state back
// to the segment start, accumulating dJ/deps along the way.  It returns
// dJ/dpsi0 and dJ/deps.  Memory: nsteps/every checkpoints plus `every`
// states of one segment.
//
// psi0 is a contiguous float64 device tensor of shape (nz, ny, nx); eps is a
// float64 scalar tensor.  mixed = true runs the FP32-mixed mode: the field is
// stepped as its FP32 deviation d = psi - m from the exact mean m (kept in
// double, conserved), and the adjoint runs in FP32 at psi = m + d; psi_T and
// the gradients are returned in float64.  Everything runs on the default
// stream, as torch's own kernels do unless the caller selects another stream.
#include <torch/extension.h>

#include "../../include/pfc_hip.h"

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

void step(pfc_hip_plan_t plan, torch::Tensor &psi, int n) {
  if (n > 0) check(pfc_hip_pfc_step(plan, psi.data_ptr(), n, nullptr), "pfc_trajectory: step");
}

void check_field(const torch::Tensor &psi, const char *what) {
  if (!psi.is_cuda() || psi.scalar_type() != torch::kFloat64 || psi.dim() != 3 ||
      !psi.is_contiguous())
    throw std::invalid_argument(std::string(what) +
                                ": psi must be a contiguous float64 device tensor (nz, ny, nx)");
}

class PfcTrajectory : public torch::autograd::Function<PfcTrajectory> {
public:
  static torch::Tensor forward(torch::autograd::AutogradContext *ctx, torch::Tensor psi0,
                               torch::Tensor eps, int64_t nsteps, double dt,
                               std::vector<double> spacing, int64_t every, bool mixed) {
    check_field(psi0, "pfc_trajectory");
    if (nsteps < 1 || every < 1) throw std::invalid_argument("pfc_trajectory: nsteps, every >= 1");
    const int nz = (int)psi0.size(0), ny = (int)psi0.size(1), nx = (int)psi0.size(2);
    pfc_hip_plan_t plan = plan_for(nx, ny, nz, mixed ? PFC_HIP_F32 : PFC_HIP_F64);
    const double e = eps.item<double>();
    const double mean = mixed ? psi0.mean().item<double>() : 0.0;
    set_ops(plan, spacing, e, dt, mixed, mean);
    std::vector<torch::Tensor> ckpts;
    torch::Tensor psi = mixed ? (psi0 - mean).to(torch::kFloat32).contiguous() : psi0.clone();
    for (int64_t s = 0; s < nsteps; s += every) {
      ckpts.push_back(psi.clone());
      step(plan, psi, (int)std::min<int64_t>(every, nsteps - s));
    }
    ctx->saved_data["ckpts"] = ckpts;
    ctx->saved_data["nsteps"] = nsteps;
    ctx->saved_data["every"] = every;
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
    const int64_t nsteps = ctx->saved_data["nsteps"].toInt();
    const int64_t every = ctx->saved_data["every"].toInt();
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
                                       nullptr),
              "pfc_trajectory: adjoint step");
    }
    torch::Tensor deps = torch::full({}, dr, eps_like.options());
    return {lam.to(torch::kFloat64), deps, torch::Tensor(), torch::Tensor(),
            torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

torch::Tensor pfc_trajectory(torch::Tensor psi0, torch::Tensor eps, int64_t nsteps, double dt,
                             std::vector<double> spacing, int64_t every, bool mixed) {
  return PfcTrajectory::apply(psi0, eps, nsteps, dt, spacing, every, mixed);
}

} // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "libtorch autograd wrapper of the MI355X PFC trajectory (forward and adjoint)";
  m.def("pfc_trajectory", &pfc_trajectory, "psi_T of nsteps PFC ETD1 steps (differentiable)",
        pybind11::arg("psi0"), pybind11::arg("eps"), pybind11::arg("nsteps"),
        pybind11::arg("dt"), pybind11::arg("spacing"), pybind11::arg("every") = 20,
        pybind11::arg("mixed") = false);
}
