// torch_recover.cpp -- the MPSFT recovery as a differentiable torch
// operator, over the C-ABI of libmpsft_hip.so:
//
//   torch.ops.mpsft.recover(x, solver, seed, stream) -> (keys, coefs)
//   torch.ops.mpsft.recover_batch(xs, batch, seeds, stream) -> (keys[], coefs[])
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

// The same over a batch (an mpsft_batch* as an integer): one batched
// recorded solve of every signal, and in the backward one batched adjoint.
class RecoverBatchFn : public torch::autograd::Function<RecoverBatchFn> {
 public:
  static variable_list forward(AutogradContext *ctx, at::TensorList xs, int64_t batch,
                               std::vector<int64_t> seeds, int64_t stream) {
    auto *b = reinterpret_cast<mpsft_batch *>(batch);
    TORCH_CHECK(b != nullptr, "mpsft.recover_batch: null batch");
    const int64_t K = int64_t(xs.size());
    TORCH_CHECK(int64_t(seeds.size()) == K, "mpsft.recover_batch: one seed per signal");
    std::vector<const void *> ptrs(K);
    std::vector<uint32_t> sd(K);
    std::vector<int64_t> ns(K);
    int32_t prec = 0;
    for (int64_t i = 0; i < K; ++i) {
      int64_t n = 0;
      Check(mpsft_batch_info(b, int32_t(i), &n, &prec, nullptr), "mpsft_batch_info");
      const auto want = prec == MPSFT_MIXED ? at::kComplexFloat : at::kComplexDouble;
      const at::Tensor &x = xs[i];
      TORCH_CHECK(x.is_cuda() && x.dim() == 1 && x.size(0) == n && x.scalar_type() == want &&
                      x.is_contiguous(),
                  "mpsft.recover_batch: signal ", i, " must be a contiguous device vector of ", n,
                  " ", prec == MPSFT_MIXED ? "complex64" : "complex128", " samples");
      ptrs[i] = x.data_ptr();
      sd[i] = uint32_t(seeds[i]);
      ns[i] = n;
    }
    void *st = reinterpret_cast<void *>(stream);
    Check(mpsft_batch_set_record(b, 1), "mpsft_batch_set_record");
    std::vector<int32_t> status(K);
    const int rc = mpsft_batch_solve(b, ptrs.data(), sd.data(), nullptr, status.data(), st);
    TORCH_CHECK(rc == MPSFT_OK || rc == MPSFT_ERR_NOT_CONVERGED || rc == MPSFT_ERR_CAPACITY,
                "mpsft: mpsft_batch_solve failed: ", mpsft_status_string(rc));
    variable_list out(2 * K);
    std::vector<int64_t> recs(K), counts(K);
    std::vector<int64_t *> kp(K);
    std::vector<void *> cp(K);
    Check(mpsft_batch_modes_counts(b, counts.data()), "mpsft_batch_modes_counts");
    for (int64_t i = 0; i < K; ++i) {
      TORCH_CHECK(status[i] == MPSFT_OK || status[i] == MPSFT_ERR_NOT_CONVERGED,
                  "mpsft.recover_batch: signal ", i, ": ", mpsft_status_string(status[i]));
      Check(mpsft_batch_info(b, int32_t(i), nullptr, nullptr, &recs[i]), "mpsft_batch_info");
      TORCH_CHECK(recs[i] != 0, "mpsft.recover_batch: signal ", i,
                  " kept no recording (its loop did not stop)");
      out[i] = at::empty({counts[i]}, xs[i].options().dtype(at::kLong));
      out[K + i] = at::empty({counts[i]}, xs[i].options().dtype(at::kComplexDouble));
      kp[i] = out[i].data_ptr<int64_t>();
      cp[i] = out[K + i].data_ptr();
    }
    Check(mpsft_batch_modes_device(b, kp.data(), cp.data(), st), "mpsft_batch_modes_device");
    ctx->saved_data["batch"] = batch;
    ctx->saved_data["stream"] = stream;
    ctx->saved_data["records"] = recs;
    ctx->saved_data["n"] = ns;
    ctx->saved_data["precision"] = int64_t(prec);
    variable_list keys(out.begin(), out.begin() + K);
    ctx->save_for_backward(keys);
    ctx->mark_non_differentiable(keys);
    return out;
  }

  static variable_list backward(AutogradContext *ctx, variable_list grads) {
    auto *b = reinterpret_cast<mpsft_batch *>(ctx->saved_data["batch"].toInt());
    const std::vector<int64_t> recs = ctx->saved_data["records"].toIntVector();
    const std::vector<int64_t> ns = ctx->saved_data["n"].toIntVector();
    const int prec = int(ctx->saved_data["precision"].toInt());
    void *st = reinterpret_cast<void *>(ctx->saved_data["stream"].toInt());
    const variable_list keys = ctx->get_saved_variables();
    const int64_t K = int64_t(keys.size());
    const auto xtype = prec == MPSFT_MIXED ? at::kComplexFloat : at::kComplexDouble;
    std::vector<at::Tensor> gs(K), outs(K);
    std::vector<const int64_t *> kp(K);
    std::vector<const void *> gp(K);
    std::vector<void *> op(K);
    std::vector<int64_t> ms(K);
    for (int64_t i = 0; i < K; ++i) {
      int64_t rec = 0;
      Check(mpsft_batch_info(b, int32_t(i), nullptr, nullptr, &rec), "mpsft_batch_info");
      TORCH_CHECK(rec == recs[i],
                  "mpsft.recover_batch backward: the batch ran another solve after this forward");
      const at::Tensor &g = grads[K + i];
      gs[i] = g.defined() ? g.to(at::kComplexDouble).contiguous()
                          : at::zeros({keys[i].numel()}, keys[i].options().dtype(at::kComplexDouble));
      outs[i] = at::empty({ns[i]}, keys[i].options().dtype(xtype));
      kp[i] = keys[i].data_ptr<int64_t>();
      gp[i] = gs[i].data_ptr();
      op[i] = outs[i].data_ptr();
      ms[i] = keys[i].numel();
    }
    Check(mpsft_batch_adjoint(b, kp.data(), gp.data(), ms.data(), op.data(), prec, st),
          "mpsft_batch_adjoint");
    variable_list res(outs.begin(), outs.end());
    res.emplace_back();  // batch
    res.emplace_back();  // seeds
    res.emplace_back();  // stream
    return res;
  }
};

std::tuple<std::vector<at::Tensor>, std::vector<at::Tensor>> RecoverBatch(
    at::TensorList xs, int64_t batch, std::vector<int64_t> seeds, int64_t stream) {
  variable_list out = RecoverBatchFn::apply(xs, batch, seeds, stream);
  const size_t K = xs.size();
  return {std::vector<at::Tensor>(out.begin(), out.begin() + K),
          std::vector<at::Tensor>(out.begin() + K, out.end())};
}

std::tuple<at::Tensor, at::Tensor> Recover(const at::Tensor &x, int64_t solver, int64_t seed,
                                           int64_t stream) {
  variable_list out = RecoverFn::apply(x, solver, seed, stream);
  return {out[0], out[1]};
}

}  // namespace

TORCH_LIBRARY(mpsft, m) {
  m.def("recover(Tensor x, int solver, int seed, int stream) -> (Tensor, Tensor)", &Recover);
  m.def("recover_batch(Tensor[] xs, int batch, int[] seeds, int stream) -> (Tensor[], Tensor[])",
        &RecoverBatch);
}
