// Drop-in host tracer: the reference's C++ interface (tritracer/tracer.h:14-126),
// same class name, constructors and method signatures, implemented over the
// C ABI of libtritracer_hip.so (include/tritracer_hip.h) instead of OptiX.
#pragma once
#include <torch/extension.h>

#include <stdexcept>
#include <string>
#include <tuple>

#include "../../include/tritracer_hip.h"

typedef torch::Tensor Tensor;

// Errors surface as Exception (a std::runtime_error), which pybind11 turns into
// a Python RuntimeError -- as the reference's exception.h:16-51.
class Exception : public std::runtime_error {
 public:
  explicit Exception(const char *msg) : std::runtime_error(msg) {}
  explicit Exception(const std::string &msg) : std::runtime_error(msg) {}
};

// Per-program state. The reference keeps the OptiX GAS, index buffer and a CUDA
// stream here (tracer.h:30-57); this tracer keeps the vertex/cumsum buffers the
// acceleration structure was last built from. Work runs on torch's current
// HIP stream.
struct TracerState {
  const float *d_vertices = nullptr;
  const int *d_cumsum = nullptr;
  Tensor indexbuf_tensor;  // keeps the cumsum the structure was built from alive
};

struct FwdTraceResult {
  std::tuple<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor> data;
  // The forward's stored blended hits (int32 [cap, H*W, 2]), counts (int32 [H*W]) and
  // colours (float [cap, H*W, 3]) for trace_bwd; undefined when the hit cache is off or the
  // program is not "forward".
  torch::Tensor hits, hit_counts, hit_colors;
};

struct BwdTraceResult {
  std::tuple<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor> data;
};

class OptixTracer {
 public:
  TracerState fwdState;
  TracerState bwdState;

  // `forward` selects the forward program: forward | reflect | refract | envlight
  // (any extension, e.g. "reflect.ptx"); `backward` must name the backward
  // program. An unknown name throws, as a missing PTX file does in the reference.
  OptixTracer(const std::string &pkgDir) : OptixTracer(pkgDir, "forward.ptx", "backward.ptx") {}
  OptixTracer(const std::string &pkgDir, const std::string &forward, const std::string &backward);

  void build_tri_accel(TracerState &state, uint32_t numVertices, uint32_t numTri, bool rebuild);

  FwdTraceResult trace_fwd(const Tensor &ray_o, const Tensor &ray_d, const Tensor &triangles_points,
                           const Tensor &sigma, const Tensor &num_points_per_triangle,
                           const Tensor &cumsum_of_points_per_triangle, const int number_of_points, const Tensor &shs,
                           const int sh_degree, const Tensor &colors_precomp, const Tensor &opacities, Tensor &scaling,
                           Tensor &density_factor, const Tensor &bg_color, const Tensor &extra,
                           const int rebuild_gas_opt, const bool debug);

  BwdTraceResult trace_bwd(const Tensor &ray_o, const Tensor &ray_d, const Tensor &background,
                           const Tensor &triangles_points, const Tensor &sigma,
                           const torch::Tensor &num_points_per_triangle,
                           const torch::Tensor &cumsum_of_points_per_triangle, const torch::Tensor &colors,
                           const torch::Tensor &others, const int number_of_points, const Tensor &shs,
                           const int degree, const Tensor &opacities, const Tensor &dL_dpixels,
                           const Tensor &dL_dothers, const Tensor &scaling, const bool debug,
                           const c10::optional<Tensor> &hits = c10::nullopt,
                           const c10::optional<Tensor> &hit_counts = c10::nullopt,
                           const c10::optional<Tensor> &hit_colors = c10::nullopt);

  // Options this tracer adds (the reference hard-codes CHUNK_SIZE 16 and a 1e-4
  // transmittance cutoff, config.h:10 and forward.cu:134). hit_cache: blended hits per
  // pixel the forward stores so that trace_bwd reads them instead of re-traversing
  // (0: re-traverse, as the reference does; same gradients either way).
  void set_options(int k, double cutoff, int chunk_mode, int hit_cache);
  // Whether the next trace_fwd stores its blended hits (default true). The Python autograd
  // Function clears it when no input needs a gradient (no backward will read them).
  void set_store_hits(bool store) { store_hits_ = store; }
  // rebuild_gas = 3 (this build's addition): refit, then rebuild when the refit tree's SAH cost
  // exceeds max_ratio x its cost right after the last build (default 1.25)
  void set_rebuild_policy(double max_ratio);
  // builds, refits, automatic rebuilds, SAH cost after the last build, last evaluated cost
  std::tuple<int, int, int, double, double> build_stats();
  std::tuple<int, double, int, int> get_options() const;
  std::string variant_name() const;

  ~OptixTracer();

  OptixTracer(OptixTracer &&other) noexcept;
  OptixTracer &operator=(OptixTracer &&other) noexcept;
  OptixTracer(const OptixTracer &) = delete;
  OptixTracer &operator=(const OptixTracer &) = delete;

 private:
  tt_tracer *tr_ = nullptr;
  int variant_ = TT_VARIANT_FORWARD;
  int k_ = 16;
  float cutoff_ = 1e-4f;
  int chunk_mode_ = TT_CHUNK_COMPAT;
  int hit_cache_ = 64;
  bool store_hits_ = true;
  int device_ = -1;
};
