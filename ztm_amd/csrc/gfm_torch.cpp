// libtorch autograd bindings of libztm_gfm (module ztm_amd._ztm_torch).
//
// GravityFn: g = G rho through ztm_plan_forward, with backward G^T (dL/dg)
//   through the hand-written adjoint kernel (ztm_plan_adjoint), so a
//   nonlinear or regularised inversion gets exact misfit gradients from
//   torch.autograd. ZTM itself differentiates by finite differences, one full
//   forward model per parameter (Jac, TFM.cpp:124-135).
// GravityBatchedFn: the same for an ensemble, Y = G P for P float32
//   [B, N] on split-BF16 or split-FP16 MFMA, as the plan's batched tuning
//   selects (ztm_plan_forward_batched), backward G^T dY through the batched
//   adjoint (ztm_plan_adjoint_batched).
// SmoothnessFn: s = 0.5 ||grad rho||^2 (forward differences), backward
//   (L^T L rho) * ds, through ztm_plan_smoothness.
//
// All run on the current HIP stream of rho's device, so they compose with
// torch work and with stream capture the same way torch's own ops do.
//
// Lifetime: each node holds a reference on its plan (ztm_plan_retain, dropped
// when autograd frees the node), so backward stays valid after the caller has
// destroyed its own plan object, e.g. a temporary Plan or a `with` block that
// has exited.
// Threshold: the forward skips cells with |rho| <= threshold (and NaN cells),
// so on a plan with threshold > 0 the gradient is 0 on those cells and G^T r
// elsewhere: the derivative of the forward as computed. With threshold 0 the
// forward is G rho for every finite rho, and only NaN cells get 0.
#include <torch/extension.h>
#include <torch/custom_class.h>
#include <c10/hip/HIPStream.h>

#include "../../include/ztm_gfm.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void check(int rc) {
    TORCH_CHECK(rc == ZTM_OK, "libztm_gfm: ", ztm_strerror(rc), ": ", ztm_last_error());
}

int precision_of(const torch::Tensor& t, int64_t n, const char* name) {
    TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor");
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
    TORCH_CHECK(t.numel() == n, name, " must have ", n, " elements, got ", t.numel());
    if (t.scalar_type() == torch::kFloat32) return ZTM_FP32;
    TORCH_CHECK(t.scalar_type() == torch::kFloat64, name, " must be float32 or float64");
    return ZTM_FP64;
}

void* stream_of(const torch::Tensor& t) {
    return (void*)c10::hip::getCurrentHIPStream(t.device().index()).stream();
}

ztm_plan as_plan(int64_t h) {
    TORCH_CHECK(h != 0, "null plan handle");
    return reinterpret_cast<ztm_plan>(h);
}

// One reference on a plan, held by an autograd node's saved data.
struct PlanRef : torch::CustomClassHolder {
    ztm_plan plan;
    double threshold = 0.0;
    explicit PlanRef(ztm_plan p) : plan(p) {
        check(ztm_plan_retain(p));
        ztm_problem prob;
        if (ztm_plan_get_problem(p, &prob) == ZTM_OK) threshold = prob.threshold;
    }
    ~PlanRef() override { ztm_plan_destroy(plan); }
};

const auto plan_ref_class = torch::class_<PlanRef>("ztm_amd", "PlanRef");

c10::intrusive_ptr<PlanRef> saved_plan(AutogradContext* ctx) {
    return ctx->saved_data["plan"].toCustomClass<PlanRef>();
}

// Zero the gradient on the cells the forward skipped (|rho| <= threshold,
// compared in rho's precision as the kernels compare it, or NaN).
torch::Tensor mask_skipped(torch::Tensor grad, const torch::Tensor& rho, double threshold) {
    auto keep = threshold > 0.0
                    ? rho.abs() > torch::scalar_tensor(threshold, rho.options())
                    : rho.isnan().logical_not();
    return torch::where(keep, grad, torch::zeros({}, grad.options()));
}

class GravityFn : public torch::autograd::Function<GravityFn> {
  public:
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor rho, int64_t plan) {
        ztm_plan p = as_plan(plan);
        const int64_t n = ztm_plan_num_cells(p), m = ztm_plan_num_stations(p);
        const int prec = precision_of(rho, n, "rho");
        auto g = torch::empty({m}, rho.options().dtype(torch::kFloat64));
        check(ztm_plan_forward(p, prec, rho.data_ptr(), g.data_ptr<double>(), stream_of(rho)));
        ctx->saved_data["plan"] = c10::make_intrusive<PlanRef>(p);
        ctx->save_for_backward({rho});
        return g;
    }

    static variable_list backward(AutogradContext* ctx, variable_list grads) {
        const auto ref = saved_plan(ctx);
        ztm_plan p = ref->plan;
        const auto rho = ctx->get_saved_variables()[0];
        const int64_t n = ztm_plan_num_cells(p), m = ztm_plan_num_stations(p);
        auto r = grads[0].to(rho.scalar_type()).contiguous();
        const int prec = precision_of(r, m, "grad_g");
        auto out = torch::empty({n}, r.options());
        check(ztm_plan_adjoint(p, prec, r.data_ptr(), out.data_ptr(), stream_of(r)));
        return {mask_skipped(out, rho, ref->threshold), torch::Tensor()};
    }
};

class GravityBatchedFn : public torch::autograd::Function<GravityBatchedFn> {
  public:
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor rho, int64_t plan) {
        ztm_plan p = as_plan(plan);
        const int64_t n = ztm_plan_num_cells(p), m = ztm_plan_num_stations(p);
        TORCH_CHECK(rho.is_cuda() && rho.is_contiguous() && rho.dim() == 2 && rho.size(1) == n &&
                        rho.scalar_type() == torch::kFloat32,
                    "rho must be a contiguous float32 GPU tensor [B, ", n, "]");
        const int64_t B = rho.size(0);
        TORCH_CHECK(B <= 256, "at most 256 models per call, got ", B);
        auto g = torch::empty({B, m}, rho.options().dtype(torch::kFloat64));
        check(ztm_plan_forward_batched(p, (int)B, rho.data_ptr<float>(), g.data_ptr<double>(),
                                       stream_of(rho)));
        ctx->saved_data["plan"] = c10::make_intrusive<PlanRef>(p);
        ctx->save_for_backward({rho});
        return g;
    }

    static variable_list backward(AutogradContext* ctx, variable_list grads) {
        const auto ref = saved_plan(ctx);
        ztm_plan p = ref->plan;
        const auto rho = ctx->get_saved_variables()[0];
        const int64_t n = ztm_plan_num_cells(p);
        auto r = grads[0].to(torch::kFloat32).contiguous();
        auto out = torch::empty({r.size(0), n}, r.options());
        check(ztm_plan_adjoint_batched(p, (int)r.size(0), r.data_ptr<float>(),
                                       out.data_ptr<float>(), stream_of(r)));
        return {mask_skipped(out, rho, ref->threshold), torch::Tensor()};
    }
};

class SmoothnessFn : public torch::autograd::Function<SmoothnessFn> {
  public:
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor rho, int64_t plan) {
        ztm_plan p = as_plan(plan);
        const int prec = precision_of(rho, ztm_plan_num_cells(p), "rho");
        auto ltl = torch::empty_like(rho);
        check(ztm_plan_smoothness(p, prec, rho.data_ptr(), ltl.data_ptr(), stream_of(rho)));
        ctx->save_for_backward({ltl});
        // 0.5 ||L rho||^2 = 0.5 rho . (L^T L rho), accumulated in FP64
        return 0.5 * torch::dot(rho.to(torch::kFloat64), ltl.to(torch::kFloat64));
    }

    static variable_list backward(AutogradContext* ctx, variable_list grads) {
        auto ltl = ctx->get_saved_variables()[0];
        return {ltl * grads[0].to(ltl.scalar_type()), torch::Tensor()};
    }
};

torch::Tensor gravity(torch::Tensor rho, int64_t plan) { return GravityFn::apply(rho, plan); }

torch::Tensor gravity_batched(torch::Tensor rho, int64_t plan) {
    return GravityBatchedFn::apply(rho, plan);
}

torch::Tensor smoothness(torch::Tensor rho, int64_t plan) {
    return SmoothnessFn::apply(rho, plan);
}

// G^T r without autograd (the adjoint on its own).
torch::Tensor adjoint(torch::Tensor r, int64_t plan) {
    ztm_plan p = as_plan(plan);
    const int prec = precision_of(r, ztm_plan_num_stations(p), "r");
    auto out = torch::empty({ztm_plan_num_cells(p)}, r.options());
    check(ztm_plan_adjoint(p, prec, r.data_ptr(), out.data_ptr(), stream_of(r)));
    return out;
}

}  // namespace

PYBIND11_MODULE(_ztm_torch, m) {
    m.doc() = "libtorch autograd bindings of libztm_gfm (gravity forward/adjoint)";
    m.def("gravity", &gravity, "g = G rho (FP64 out); differentiable, backward = G^T grad",
          pybind11::arg("rho"), pybind11::arg("plan"));
    m.def("gravity_batched", &gravity_batched,
          "Y = G P for P float32 [B, N] (FP64 out) on split-BF16/FP16 MFMA (the plan's "
          "batched tuning); differentiable, backward = batched G^T grad",
          pybind11::arg("rho"), pybind11::arg("plan"));
    m.def("smoothness", &smoothness, "0.5 ||grad rho||^2; differentiable",
          pybind11::arg("rho"), pybind11::arg("plan"));
    m.def("adjoint", &adjoint, "G^T r", pybind11::arg("r"), pybind11::arg("plan"));
}
