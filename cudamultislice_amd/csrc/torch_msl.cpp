// torch_msl.cpp -- thin torch extension over the differentiable path of
// libmsl (include/msl.h, msl_diff_*): a TorchScript custom class
// torch.classes.msl.DiffEngine whose forward / backward take and return
// tensors and call the C-ABI on their device pointers.  All arithmetic runs
// in libmsl's HIP kernels; torch supplies the device memory and the stream
// (passed as the hipStream_t handle torch.cuda.current_stream().cuda_stream).
#include <ATen/ATen.h>
#include <torch/custom_class.h>
#include <torch/library.h>

#include <string>
#include <vector>

#include "../../include/msl.h"

namespace {

struct DiffEngine : torch::CustomClassHolder {
  msl_diff* h = nullptr;
  int64_t nx, ny, nslices, nprobe;

  DiffEngine(int64_t device, int64_t nx_, int64_t ny_, double ps, double kV, double dz, int64_t nslices_,
             int64_t nprobe_, int64_t checkpoint_every)
      : nx(nx_), ny(ny_), nslices(nslices_), nprobe(nprobe_) {
    const msl_status st = msl_diff_create(&h, (int)device, (int)nx, (int)ny, (float)ps, (float)kV, (float)dz,
                                          (int)nslices, (int)nprobe, (int)checkpoint_every);
    TORCH_CHECK(st == MSL_OK, "msl status ", st, " (", msl_status_string(st), "): msl_diff_create");
  }
  ~DiffEngine() override { msl_diff_destroy(h); }

  void check(msl_status st) const {
    TORCH_CHECK(st == MSL_OK, "msl status ", st, " (", msl_status_string(st), "): ", msl_diff_last_error(h));
  }

  // |psi_exit|^2 for potentials V [nslices][ny][nx] (float32) and entrance
  // waves psi0 [nprobe][ny][nx] (complex64); keep = false keeps nothing for a
  // backward (msl_diff_forward_only)
  at::Tensor forward(const at::Tensor& V, const at::Tensor& psi0, bool keep, int64_t stream) {
    TORCH_CHECK(V.is_cuda() && psi0.is_cuda(), "V and psi0 must be on the GPU");
    TORCH_CHECK(V.scalar_type() == at::kFloat && V.is_contiguous() && V.dim() == 3 && V.size(0) == nslices &&
                    V.size(1) == ny && V.size(2) == nx,
                "V must be contiguous float32 [nslices, ny, nx]");
    TORCH_CHECK(psi0.scalar_type() == at::kComplexFloat && psi0.is_contiguous() && psi0.dim() == 3 &&
                    psi0.size(0) == nprobe && psi0.size(1) == ny && psi0.size(2) == nx,
                "psi0 must be contiguous complex64 [nprobe, ny, nx]");
    at::Tensor out = at::empty({nprobe, ny, nx}, V.options());
    void* s = reinterpret_cast<void*>(stream);
    check(keep ? msl_diff_forward(h, V.data_ptr<float>(), psi0.data_ptr(), out.data_ptr<float>(), s)
               : msl_diff_forward_only(h, V.data_ptr<float>(), psi0.data_ptr(), out.data_ptr<float>(), s));
    return out;
  }

  // dL/dV [nslices][ny][nx] from dL/dI [nprobe][ny][nx] after a keeping forward
  at::Tensor backward(const at::Tensor& dLdI, int64_t stream) {
    TORCH_CHECK(dLdI.is_cuda() && dLdI.dim() == 3 && dLdI.size(0) == nprobe && dLdI.size(1) == ny &&
                    dLdI.size(2) == nx,
                "dL/dI must be a [nprobe, ny, nx] GPU tensor");
    const at::Tensor g = dLdI.to(at::kFloat).contiguous();
    at::Tensor out = at::empty({nslices, ny, nx}, g.options());
    check(msl_diff_backward(h, g.data_ptr<float>(), out.data_ptr<float>(), reinterpret_cast<void*>(stream)));
    return out;
  }

  void set_chunk(int64_t probes) { check(msl_diff_set_chunk(h, (int)probes)); }

  // wavelength, sigma, kmax, workspace bytes
  std::vector<double> info() const {
    float wl = 0, sg = 0, km = 0;
    long long ws = 0;
    check(msl_diff_info(h, &wl, &sg, &km, &ws));
    return {wl, sg, km, (double)ws};
  }
};

}  // namespace

TORCH_LIBRARY(msl, m) {
  m.class_<DiffEngine>("DiffEngine")
      .def(torch::init<int64_t, int64_t, int64_t, double, double, double, int64_t, int64_t, int64_t>())
      .def("forward", &DiffEngine::forward)
      .def("backward", &DiffEngine::backward)
      .def("set_chunk", &DiffEngine::set_chunk)
      .def("info", &DiffEngine::info);
}
