// bimp_autograd.cpp -- libtorch autograd Function over the differentiation form's C ABI
// (bimp_diff_*, include/bimp_hip.h).  The forward maps a batch of float images to the
// per-scale response maps D; the backward is the library's hand-written adjoint, giving the
// gradient with respect to the images and to the Gabor parameters (sigma ratio, gamma).
// Work runs on the HIP stream handle the caller passes (the current stream of the forward).
#include <torch/extension.h>

#include "bimp_hip.h"

namespace {

void check(int st, const bimp_diff* d) {
    TORCH_CHECK(st >= 0, "bimp: ", bimp_diff_last_error(d), " (", bimp_status_string(st), ")");
}

void set_params(bimp_diff* d, const torch::Tensor& params) {
    TORCH_CHECK(params.numel() == 2, "params must hold (sigma_ratio, gamma)");
    const auto p = params.detach().to(torch::kCPU, torch::kFloat64).contiguous();
    const double* v = p.data_ptr<double>();
    check(bimp_diff_set_params(d, (float)v[0], (float)v[1]), d);
}

class BimpResponse : public torch::autograd::Function<BimpResponse> {
   public:
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor image,
                                 torch::Tensor params, int64_t handle, int64_t stream) {
        auto* d = reinterpret_cast<bimp_diff*>(handle);
        TORCH_CHECK(image.is_cuda() && image.scalar_type() == torch::kFloat32 && image.dim() == 3,
                    "image must be a (N, H, W) float32 device tensor");
        auto img = image.contiguous();
        set_params(d, params);
        auto resp = torch::empty({img.size(0), (int64_t)bimp_diff_response_size(d)}, img.options());
        check(bimp_diff_forward(d, img.data_ptr<float>(), (int)img.size(0), resp.data_ptr<float>(),
                                reinterpret_cast<bimp_stream_t>(stream)),
              d);
        ctx->save_for_backward({img, params});
        ctx->saved_data["handle"] = handle;
        ctx->saved_data["stream"] = stream;
        return resp;
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grads) {
        auto saved = ctx->get_saved_variables();
        auto img = saved[0], params = saved[1];
        auto* d = reinterpret_cast<bimp_diff*>(ctx->saved_data["handle"].toInt());
        const int64_t stream = ctx->saved_data["stream"].toInt();
        auto g = grads[0].contiguous();
        set_params(d, params);
        auto gimg = torch::empty_like(img);
        double gp[2] = {0, 0};
        check(bimp_diff_backward(d, img.data_ptr<float>(), (int)img.size(0), g.data_ptr<float>(),
                                 gimg.data_ptr<float>(), gp, reinterpret_cast<bimp_stream_t>(stream)),
              d);
        auto gparams = torch::tensor({gp[0], gp[1]}, torch::kFloat64).to(params.options());
        return {gimg, gparams, torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor response(torch::Tensor image, torch::Tensor params, int64_t handle, int64_t stream) {
    return BimpResponse::apply(image, params, handle, stream);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("response", &response, "BIMP response maps with the hand-written adjoint",
          pybind11::arg("image"), pybind11::arg("params"), pybind11::arg("handle"),
          pybind11::arg("stream"));
}
