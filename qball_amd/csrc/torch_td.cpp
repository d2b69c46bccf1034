// libtorch autograd wrapper of the TD objective: J(v_ext, A0) with
//   A_n = A0 * shape_n (the field amplitude times its fixed time profile),
//   J = sum_n w_n e.j(psi_n, A_n)   (qbh_td_forward),
// and gradients dJ/dv_ext, dJ/dA0 from the hand-written adjoint (qbh_td_backward).
// The backward re-runs the forward from the saved inputs first, so a forward call made on
// the same propagator between this node's forward and backward cannot change the result;
// checkpoints every `checkpoint` steps bound the memory of both passes.
#include <torch/cuda.h>
#include <torch/extension.h>

#include <string>

#include "../../include/qbh.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::Function;
using torch::autograd::variable_list;

struct Args {
  qbh_td td;
  const double* psi0;
  long ld;
  int nsteps;
  double dt;
  int order, checkpoint, flags;
};

double run_forward(const Args& a, const torch::Tensor& vext, double A0, const torch::Tensor& shape,
                   const torch::Tensor& w, const torch::Tensor& e) {
  // The propagator runs on its context's stream, which need not be torch's current one:
  // wait for the device so that vext (produced on any torch stream) is complete before the
  // context reads it.  qbh_td_forward / qbh_td_backward synchronise their own stream before
  // returning, so their outputs are ready for torch afterwards.
  torch::cuda::synchronize(vext.device().index());
  torch::Tensor A = (shape * A0).contiguous();
  double J = 0.0;
  const int r = qbh_td_forward(a.td, a.psi0, a.ld, a.nsteps, a.dt, a.order,
                               A.data_ptr<double>(), w.data_ptr<double>(), e.data_ptr<double>(),
                               vext.defined() ? vext.data_ptr<double>() : nullptr, a.checkpoint,
                               a.flags, &J, nullptr, 0);
  TORCH_CHECK(r == QBH_OK, "qbh_td_forward failed (", r, ")");
  return J;
}

struct TDObjective : public Function<TDObjective> {
  static torch::Tensor forward(AutogradContext* ctx, torch::Tensor vext, torch::Tensor A0,
                               int64_t td, int64_t psi0, int64_t ld, int64_t nsteps, double dt,
                               torch::Tensor shape, torch::Tensor w, torch::Tensor e,
                               int64_t order, int64_t checkpoint, int64_t flags) {
    TORCH_CHECK(vext.is_cuda() && vext.scalar_type() == torch::kFloat64 && vext.is_contiguous(),
                "vext: contiguous float64 CUDA tensor");
    TORCH_CHECK(!shape.is_cuda() && shape.scalar_type() == torch::kFloat64 &&
                    shape.numel() == 3 * (nsteps + 1),
                "shape: float64 host tensor (nsteps+1, 3)");
    TORCH_CHECK(!w.is_cuda() && w.numel() == nsteps + 1, "w: float64 host tensor (nsteps+1)");
    TORCH_CHECK(!e.is_cuda() && e.numel() == 3, "e: float64 host tensor (3)");
    Args a{reinterpret_cast<qbh_td>(td), reinterpret_cast<const double*>(psi0), (long)ld,
           (int)nsteps, dt, (int)order, (int)checkpoint, (int)flags};
    const double J = run_forward(a, vext, A0.item<double>(), shape.contiguous(), w.contiguous(),
                                 e.contiguous());
    ctx->save_for_backward({vext, A0, shape.contiguous(), w.contiguous(), e.contiguous()});
    ctx->saved_data["td"] = td;
    ctx->saved_data["psi0"] = psi0;
    ctx->saved_data["ld"] = ld;
    ctx->saved_data["nsteps"] = nsteps;
    ctx->saved_data["dt"] = dt;
    ctx->saved_data["order"] = order;
    ctx->saved_data["checkpoint"] = checkpoint;
    ctx->saved_data["flags"] = flags;
    return torch::full({}, J, torch::dtype(torch::kFloat64));
  }

  static variable_list backward(AutogradContext* ctx, variable_list grad_out) {
    auto saved = ctx->get_saved_variables();
    torch::Tensor vext = saved[0], A0 = saved[1], shape = saved[2], w = saved[3], e = saved[4];
    Args a{reinterpret_cast<qbh_td>(ctx->saved_data["td"].toInt()),
           reinterpret_cast<const double*>(ctx->saved_data["psi0"].toInt()),
           (long)ctx->saved_data["ld"].toInt(), (int)ctx->saved_data["nsteps"].toInt(),
           ctx->saved_data["dt"].toDouble(), (int)ctx->saved_data["order"].toInt(),
           (int)ctx->saved_data["checkpoint"].toInt(), (int)ctx->saved_data["flags"].toInt()};
    run_forward(a, vext, A0.item<double>(), shape, w, e);
    torch::Tensor gv = torch::empty_like(vext);
    torch::Tensor gA = torch::zeros({a.nsteps + 1, 3}, torch::dtype(torch::kFloat64));
    const int r = qbh_td_backward(a.td, gv.data_ptr<double>(), gA.data_ptr<double>());
    TORCH_CHECK(r == QBH_OK, "qbh_td_backward failed (", r, ")");
    const double g = grad_out[0].item<double>();
    const double dA0 = (gA * shape.view({a.nsteps + 1, 3})).sum().item<double>();
    torch::Tensor gA0 = torch::full({}, g * dA0, A0.options());
    return {gv * g,        gA0,           torch::Tensor(), torch::Tensor(), torch::Tensor(),
            torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(),
            torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

torch::Tensor td_objective(torch::Tensor vext, torch::Tensor A0, int64_t td, int64_t psi0,
                           int64_t ld, int64_t nsteps, double dt, torch::Tensor shape,
                           torch::Tensor w, torch::Tensor e, int64_t order, int64_t checkpoint,
                           int64_t flags) {
  return TDObjective::apply(vext, A0, td, psi0, ld, nsteps, dt, shape, w, e, order, checkpoint,
                            flags);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "autograd Function over the qbh TD propagator and its adjoint";
  m.def("td_objective", &td_objective,
        "J(v_ext, A0) of qbh_td_forward with dJ/dv_ext and dJ/dA0 from qbh_td_backward");
}
