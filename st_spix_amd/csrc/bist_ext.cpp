// Thin torch C++ extension: the reference's bist_cuda module
// (csrc/pybind_bist.cpp:16-17, init_bist in csrc/bist/pyapi.cu:307-309),
// bist_forward (pyapi.cu:196-215, main_loop 46-195) and get_marked_video
// (pyapi.cu:260-303), on libstspix.so's C ABI.  Host code only: it checks the
// tensors, allocates the outputs with torch's caching allocator on the
// tensors' device and launches on torch's current HIP stream.
#include <torch/extension.h>

// PyTorch-ROCm names HIP devices "cuda": its guard and stream types for such
// tensors are the MasqueradingAsCUDA ones
#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>

#include <vector>

#include "../../include/stspix.h"

namespace {

#define STSP_CHECK_INPUT(x)                                               \
  TORCH_CHECK((x).is_cuda(), #x " must be a CUDA tensor");               \
  TORCH_CHECK((x).is_contiguous(), #x " must be contiguous")

void check_status(int rc, const char* where) {
  TORCH_CHECK(rc == STSP_OK, where, ": ", stsp_status_string(rc));
}

// The parameter arrays of one frame (the reference SuperpixelParams fields),
// cap entries each, as tensors; `st` points at them.
struct FrameParams {
  std::vector<torch::Tensor> keep;
  stsp_params st{};
  explicit FrameParams(int cap, const torch::Device& dev) {
    auto f32 = torch::TensorOptions().dtype(torch::kFloat32).device(dev);
    auto f64 = torch::TensorOptions().dtype(torch::kFloat64).device(dev);
    auto i32 = torch::TensorOptions().dtype(torch::kInt32).device(dev);
    auto make = [&](int64_t n, const torch::TensorOptions& o) {
      keep.push_back(torch::empty({n}, o));
      return keep.back().data_ptr();
    };
    st.mu_app = static_cast<float*>(make(3ll * cap, f32));
    st.prior_mu_app = static_cast<float*>(make(3ll * cap, f32));
    st.mu_shape = static_cast<double*>(make(2ll * cap, f64));
    st.sigma_shape = static_cast<double*>(make(3ll * cap, f64));
    st.logdet_sigma_shape = static_cast<float*>(make(cap, f32));
    st.prior_mu_shape = static_cast<double*>(make(2ll * cap, f64));
    st.prior_sigma_shape = static_cast<double*>(make(3ll * cap, f64));
    st.sample_sigma_shape = static_cast<double*>(make(3ll * cap, f64));
    st.counts = static_cast<int*>(make(cap, i32));
    st.sm_counts = static_cast<int*>(make(cap, i32));
    st.prior_counts = static_cast<float*>(make(cap, f32));
    st.ids = static_cast<int*>(make(cap, i32));
    st.invalid_counts = static_cast<float*>(make(cap, f32));
    st.cap = cap;
  }
};

struct Ctx {
  stsp_ctx* c = nullptr;
  ~Ctx() {
    if (c) stsp_ctx_destroy(c);
  }
};

// bist_forward_cuda(vid, flows, niters, sp_size, potts, sigma2_app, alpha,
// split_alpha, video_mode, rgb2lab): labels (T, H, W) int32 of a video
// (T, H, W, 3) float32 with flows (>= T-1, H, W, 2) float32, flows[t-1]
// the flow into frame t.  rgb2lab is accepted and unused, as in the
// reference (pyapi.cu:97-101).
torch::Tensor bist_forward(const torch::Tensor& vid, const torch::Tensor& flows, int niters,
                           int sp_size, double potts, double sigma2_app, double alpha,
                           double split_alpha, bool video_mode, bool rgb2lab) {
  (void)rgb2lab;
  STSP_CHECK_INPUT(vid);
  STSP_CHECK_INPUT(flows);
  TORCH_CHECK(vid.scalar_type() == torch::kFloat32 && vid.dim() == 4 && vid.size(3) == 3,
              "vid must be float32 (T, H, W, 3)");
  const int64_t T = vid.size(0), H = vid.size(1), W = vid.size(2);
  TORCH_CHECK(T >= 1 && H >= 1 && W >= 1, "vid must not be empty");
  if (video_mode && T > 1) {
    TORCH_CHECK(flows.scalar_type() == torch::kFloat32 && flows.dim() == 4 &&
                    flows.size(0) >= T - 1 && flows.size(1) == H && flows.size(2) == W &&
                    flows.size(3) == 2,
                "flows must be float32 (T-1, H, W, 2)");
    TORCH_CHECK(flows.device() == vid.device(), "flows must be on vid's device");
  }
  const c10::hip::HIPGuardMasqueradingAsCUDA guard(vid.device());
  hipStream_t stream = c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(vid.device().index()).stream();
  stsp_config cfg;
  stsp_default_config(&cfg);
  cfg.niters = niters;
  cfg.sp_size = sp_size;
  cfg.potts = (float)potts;
  cfg.sigma2_app = (float)sigma2_app;
  cfg.alpha_hastings = (float)alpha;
  cfg.split_alpha = (float)split_alpha;
  Ctx ctx;
  check_status(stsp_ctx_create(&cfg, (int)H, (int)W, &ctx.c), "stsp_ctx_create");
  auto out = torch::empty({T, H, W}, vid.options().dtype(torch::kInt32));
  std::unique_ptr<FrameParams> prev;
  int nprev = 0;
  const int64_t npix = H * W;
  for (int64_t t = 0; t < T; t++) {
    const float* img = vid.data_ptr<float>() + t * npix * 3;
    int* spix = out.data_ptr<int>() + t * npix;
    int nout = 0;
    if (t == 0 || !video_mode) {
      auto cur = std::make_unique<FrameParams>(stsp_bass_capacity(sp_size, (int)H, (int)W),
                                               vid.device());
      check_status(stsp_bass(ctx.c, img, spix, &cur->st, &nout, stream), "stsp_bass");
      prev = std::move(cur);
    } else {
      auto cur = std::make_unique<FrameParams>(stsp_prop_capacity(nprev), vid.device());
      const float* flow = flows.data_ptr<float>() + (t - 1) * npix * 2;
      const int* sprev = out.data_ptr<int>() + (t - 1) * npix;
      int niters_state = 0;
      check_status(stsp_prop(ctx.c, img, flow, sprev, &prev->st, nprev, spix, &cur->st, &nout,
                             &niters_state, stream),
                   "stsp_prop");
      prev = std::move(cur);
    }
    nprev = nout;
  }
  return out;
}

// get_marked_video(vid, spix, color): vid (T, H, W, 3) float32, spix
// (T, H, W) int32, color 3 floats; border pixels of the labels take the
// colour, the others the image clamped to [0, 1].
torch::Tensor get_marked_video(const torch::Tensor& vid, const torch::Tensor& spix,
                               const torch::Tensor& color) {
  STSP_CHECK_INPUT(vid);
  STSP_CHECK_INPUT(spix);
  TORCH_CHECK(color.is_contiguous(), "color must be contiguous");
  TORCH_CHECK(vid.scalar_type() == torch::kFloat32 && vid.dim() == 4 && vid.size(3) == 3,
              "vid must be float32 (T, H, W, 3)");
  TORCH_CHECK(spix.scalar_type() == torch::kInt32 && spix.dim() == 3 &&
                  spix.size(0) == vid.size(0) && spix.size(1) == vid.size(1) &&
                  spix.size(2) == vid.size(2),
              "spix must be int32 (T, H, W) matching vid");
  TORCH_CHECK(color.numel() == 3, "color must hold 3 values");
  const c10::hip::HIPGuardMasqueradingAsCUDA guard(vid.device());
  hipStream_t stream = c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(vid.device().index()).stream();
  auto col = color.to(torch::kCPU, torch::kFloat32).contiguous();
  auto marked = torch::empty_like(vid);
  check_status(stsp_marked_video(vid.data_ptr<float>(), spix.data_ptr<int>(),
                                 col.data_ptr<float>(), (int)vid.size(0), (int)vid.size(1),
                                 (int)vid.size(2), marked.data_ptr<float>(), stream),
               "stsp_marked_video");
  return marked;
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.def("bist_forward", &bist_forward, "BIST", py::arg("vid"), py::arg("flows"),
        py::arg("niters"), py::arg("sp_size"), py::arg("potts"), py::arg("sigma2_app"),
        py::arg("alpha"), py::arg("split_alpha"), py::arg("video_mode"),
        py::arg("rgb2lab") = false);
  m.def("get_marked_video", &get_marked_video, "get marked video");
}
