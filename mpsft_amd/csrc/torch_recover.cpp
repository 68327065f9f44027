// torch_recover.cpp -- the MPSFT recovery as a differentiable torch
// operator, over the C-ABI of libmpsft_hip.so:
//
//   torch.ops.mpsft.recover(x, solver, seed, stream) -> (keys, coefs)
//
// Forward: Demo1::Run (src/demo1.cc:36-59) on the device signal x, with the
// solve's decisions recorded (its transforms, q's, bins per iteration and
// the k1 each bin identified).  With those decisions held, Iterate
// (src/iterate.cc:80-177) is linear in x, so the recovered coefficients are
// c = F x.  keys (int64) are the found frequencies k0, coefs (complex128)
// their coefficients, both on x's device; keys carry no gradient.
// Backward: x.grad = F^H grad_coefs through mpsft_solver_adjoint.  torch's
// gradient of a real loss L with respect to a complex tensor z is
// dL/dRe z + i dL/dIm z, under which the gradient of x for c = F x is F^H of
// the gradient of c.
//
// No reference counterpart: the reference recovers, it does not
// differentiate.  The solver (an mpsft_solver* passed as an integer) must
// not run another solve between a forward and its backward; the backward
// checks the recording it was given and refuses a replaced one.  stream is
// the hipStream_t (as an integer, 0 = the null stream) of both passes.
#include <ATen/ATen.h>
#include <torch/csrc/autograd/custom_function.h>
#include <torch/library.h>

#include <tuple>

#include "mpsft.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void Check(int st, const char *what) {
  TORCH_CHECK(st == MPSFT_OK, "mpsft: ", what, " failed: ", mpsft_status_string(st));
}

class RecoverFn : public torch::autograd::Function<RecoverFn> {
 public:
  static variable_list forward(AutogradContext *ctx, const at::Tensor &x, int64_t solver,
                               int64_t seed, int64_t stream) {
    auto *s = reinterpret_cast<mpsft_solver *>(solver);
    TORCH_CHECK(s != nullptr, "mpsft.recover: null solver");
    int64_t n = 0;
    int32_t prec = 0;
    Check(mpsft_solver_info(s, &n, &prec, nullptr), "mpsft_solver_info");
    const auto want = prec == MPSFT_MIXED ? at::kComplexFloat : at::kComplexDouble;
    TORCH_CHECK(x.is_cuda() && x.dim() == 1 && x.size(0) == n && x.scalar_type() == want &&
                    x.is_contiguous(),
                "mpsft.recover: x must be a contiguous device vector of ", n, " ",
                prec == MPSFT_MIXED ? "complex64" : "complex128", " samples");
    void *st = reinterpret_cast<void *>(stream);
    Check(mpsft_solver_set_record(s, 1), "mpsft_solver_set_record");
    int32_t iters = 0;
    const int rc = mpsft_solve(s, x.data_ptr(), uint32_t(seed), &iters, st);
    TORCH_CHECK(rc == MPSFT_OK || rc == MPSFT_ERR_NOT_CONVERGED, "mpsft: mpsft_solve failed: ",
                mpsft_status_string(rc));
    int64_t m = 0;
    Check(mpsft_solver_modes_count(s, &m), "mpsft_solver_modes_count");
    at::Tensor hk = at::empty({m}, at::kLong);
    at::Tensor hc = at::empty({m}, at::kComplexDouble);
    Check(mpsft_solver_modes_get(s, hk.data_ptr<int64_t>(), reinterpret_cast<double *>(hc.data_ptr()),
                                 m, &m),
          "mpsft_solver_modes_get");
    int64_t rec = 0;
    Check(mpsft_solver_info(s, nullptr, nullptr, &rec), "mpsft_solver_info");
    TORCH_CHECK(rec != 0, "mpsft.recover: the solve kept no recording (its loop did not stop)");
    at::Tensor keys = hk.to(x.device());
    at::Tensor coefs = hc.to(x.device());
    ctx->saved_data["solver"] = solver;
    ctx->saved_data["stream"] = stream;
    ctx->saved_data["record"] = rec;
    ctx->saved_data["n"] = n;
    ctx->saved_data["precision"] = int64_t(prec);
    ctx->save_for_backward({keys});
    ctx->mark_non_differentiable({keys});
    return {keys, coefs};
  }

  static variable_list backward(AutogradContext *ctx, variable_list grads) {
    auto *s = reinterpret_cast<mpsft_solver *>(ctx->saved_data["solver"].toInt());
    const int64_t n = ctx->saved_data["n"].toInt();
    const int prec = int(ctx->saved_data["precision"].toInt());
    void *st = reinterpret_cast<void *>(ctx->saved_data["stream"].toInt());
    int64_t rec = 0;
    Check(mpsft_solver_info(s, nullptr, nullptr, &rec), "mpsft_solver_info");
    TORCH_CHECK(rec == ctx->saved_data["record"].toInt(),
                "mpsft.recover backward: the solver ran another solve after this forward");
    const at::Tensor keys = ctx->get_saved_variables()[0];
    const auto xtype = prec == MPSFT_MIXED ? at::kComplexFloat : at::kComplexDouble;
    at::Tensor gx = at::empty({n}, keys.options().dtype(xtype));
    at::Tensor g = grads[1];
    if (!g.defined()) {
      gx.zero_();
    } else {
      g = g.to(at::kComplexDouble).contiguous();
      Check(mpsft_solver_adjoint(s, keys.data_ptr<int64_t>(), g.data_ptr(), keys.numel(),
                                 gx.data_ptr(), prec, st),
            "mpsft_solver_adjoint");
    }
    return {gx, at::Tensor(), at::Tensor(), at::Tensor()};
  }
};

std::tuple<at::Tensor, at::Tensor> Recover(const at::Tensor &x, int64_t solver, int64_t seed,
                                           int64_t stream) {
  variable_list out = RecoverFn::apply(x, solver, seed, stream);
  return {out[0], out[1]};
}

}  // namespace

TORCH_LIBRARY(mpsft, m) {
  m.def("recover(Tensor x, int solver, int seed, int stream) -> (Tensor, Tensor)", &Recover);
}
