// OptixTracer over libtritracer_hip.so. Follows the reference host code
// (tritracer/tracer.cpp:16-228, pipeline.cpp:186-326) for argument meaning,
// output shapes and the rebuild_gas codes; the launches go through the C ABI.
#include "tracer.h"

#include <c10/hip/HIPGuard.h>
#include <c10/hip/HIPStream.h>
#include <hip/hip_runtime_api.h>

#include <algorithm>

namespace {

void check_tt(int rc, const char *what) {
  if (rc != TT_OK) throw Exception(std::string(what) + " failed: " + tt_last_error());
}

std::string stem(const std::string &name) {
  std::string base = name;
  size_t slash = base.find_last_of("/\\");
  if (slash != std::string::npos) base = base.substr(slash + 1);
  size_t dot = base.find('.');
  if (dot != std::string::npos) base = base.substr(0, dot);
  return base;
}

// The tensor as contiguous data of the given dtype on a CUDA (HIP) device; the
// returned tensor owns the memory for the duration of the call (the reference
// takes .contiguous().data_ptr() of a temporary, tracer.cpp:46).
Tensor prep(const Tensor &t, c10::ScalarType dt, const char *name, int device) {
  if (!t.defined()) throw Exception(std::string(name) + " is undefined");
  if (!t.is_cuda()) throw Exception(std::string(name) + " must be a CUDA (HIP) tensor");
  if (t.get_device() != device)
    throw Exception(std::string(name) + " is on device " + std::to_string(t.get_device()) + ", expected " +
                    std::to_string(device));
  if (t.scalar_type() != dt)
    throw Exception(std::string(name) + " has dtype " + std::string(c10::toString(t.scalar_type())) + ", expected " +
                    std::string(c10::toString(dt)));
  return t.contiguous();
}

void *cur_stream(int device) {
  return static_cast<void *>(c10::hip::getCurrentHIPStream(device).stream());
}

}  // namespace

OptixTracer::OptixTracer(const std::string &pkgDir, const std::string &forward, const std::string &backward) {
  (void)pkgDir;  // programs are compiled into libtritracer_hip.so; the names select them
  const std::string f = stem(forward), b = stem(backward);
  if (f == "forward")
    variant_ = TT_VARIANT_FORWARD;
  else if (f == "reflect")
    variant_ = TT_VARIANT_REFLECT;
  else if (f == "refract")
    variant_ = TT_VARIANT_REFRACT;
  else if (f == "envlight")
    variant_ = TT_VARIANT_ENVLIGHT;
  else
    throw Exception("Unknown forward program: " + forward + " (forward, reflect, refract or envlight)");
  if (b != "backward") throw Exception("Unknown backward program: " + backward + " (backward)");
  if (hipGetDevice(&device_) != hipSuccess) throw Exception("hipGetDevice failed");
  check_tt(tt_create(&tr_), "tt_create");
}

OptixTracer::~OptixTracer() {
  if (tr_) tt_destroy(tr_);
  tr_ = nullptr;
}

OptixTracer::OptixTracer(OptixTracer &&o) noexcept
    : fwdState(std::move(o.fwdState)), bwdState(std::move(o.bwdState)), tr_(o.tr_), variant_(o.variant_), k_(o.k_),
      cutoff_(o.cutoff_), chunk_mode_(o.chunk_mode_), hit_cache_(o.hit_cache_), store_hits_(o.store_hits_),
      device_(o.device_) {
  o.tr_ = nullptr;
}

OptixTracer &OptixTracer::operator=(OptixTracer &&o) noexcept {
  if (this != &o) {
    if (tr_) tt_destroy(tr_);
    fwdState = std::move(o.fwdState);
    bwdState = std::move(o.bwdState);
    tr_ = o.tr_;
    variant_ = o.variant_;
    k_ = o.k_;
    cutoff_ = o.cutoff_;
    chunk_mode_ = o.chunk_mode_;
    hit_cache_ = o.hit_cache_;
    store_hits_ = o.store_hits_;
    device_ = o.device_;
    o.tr_ = nullptr;
  }
  return *this;
}

void OptixTracer::set_options(int k, double cutoff, int chunk_mode, int hit_cache) {
  if (k != 8 && k != 16 && k != 32) throw Exception("k must be 8, 16 or 32");
  if (!(cutoff >= 0.0 && cutoff < 1.0)) throw Exception("cutoff must be in [0, 1)");
  if (chunk_mode != TT_CHUNK_COMPAT && chunk_mode != TT_CHUNK_EXACT) throw Exception("chunk_mode must be 0 or 1");
  if (hit_cache < 0 || hit_cache > 1024) throw Exception("hit_cache must be in [0, 1024]");
  k_ = k;
  cutoff_ = (float)cutoff;
  chunk_mode_ = chunk_mode;
  hit_cache_ = hit_cache;
}

std::tuple<int, double, int, int> OptixTracer::get_options() const {
  return std::make_tuple(k_, (double)cutoff_, chunk_mode_, hit_cache_);
}

std::string OptixTracer::variant_name() const {
  static const char *names[] = {"forward", "reflect", "refract", "envlight"};
  return names[variant_];
}

// pipeline.cpp:206-265. `state` supplies the vertex and cumsum buffers.
void OptixTracer::set_rebuild_policy(double max_ratio) {
  check_tt(tt_set_rebuild_policy(tr_, (float)max_ratio), "set_rebuild_policy");
}

std::tuple<int, int, int, double, double> OptixTracer::build_stats() {
  int n[3];
  float c[2];
  check_tt(tt_build_stats(tr_, n, c), "build_stats");
  return std::make_tuple(n[0], n[1], n[2], (double)c[0], (double)c[1]);
}

void OptixTracer::build_tri_accel(TracerState &state, uint32_t numVertices, uint32_t numTri, bool rebuild) {
  check_tt(tt_build(tr_, state.d_vertices, (int)numVertices, state.d_cumsum, (int)numTri,
                    rebuild ? TT_BUILD : TT_REFIT, cur_stream(device_)),
           rebuild ? "acceleration structure build" : "acceleration structure refit");
}

FwdTraceResult OptixTracer::trace_fwd(const Tensor &ray_o, const Tensor &ray_d, const Tensor &triangles_points,
                                      const Tensor &sigma, const Tensor &num_points_per_triangle,
                                      const Tensor &cumsum_of_points_per_triangle, const int number_of_points,
                                      const Tensor &shs, const int sh_degree, const Tensor &colors_precomp,
                                      const Tensor &opacities, Tensor &scaling, Tensor &density_factor,
                                      const Tensor &bg_color, const Tensor &extra, const int rebuild_gas_opt,
                                      const bool debug) {
  (void)colors_precomp;  // accepted and ignored, as in the reference (params.cols is never set)
  const int dev = device_;
  c10::hip::HIPGuard guard(dev);
  if (ray_o.dim() != 3 || ray_o.size(2) != 3) throw Exception("ray_o must be [H, W, 3]");
  if (!ray_d.sizes().equals(ray_o.sizes())) throw Exception("ray_d must have the shape of ray_o");
  const int H = (int)ray_o.size(0), W = (int)ray_o.size(1);
  const int P = number_of_points;
  if (P < 0) throw Exception("number_of_points must be non-negative");
  if (cumsum_of_points_per_triangle.numel() != P || num_points_per_triangle.numel() != P)
    throw Exception("num_points_per_triangle / cumsum_of_points_per_triangle must have number_of_points entries");

  Tensor ro = prep(ray_o, torch::kFloat32, "ray_o", dev);
  Tensor rd = prep(ray_d, torch::kFloat32, "ray_d", dev);
  Tensor verts = prep(triangles_points, torch::kFloat32, "triangles_points", dev);
  Tensor npts = prep(num_points_per_triangle, torch::kInt32, "num_points_per_triangle", dev);
  Tensor cums = prep(cumsum_of_points_per_triangle, torch::kInt32, "cumsum_of_points_per_triangle", dev);
  Tensor sg = prep(sigma, torch::kFloat32, "sigma", dev);
  Tensor op = prep(opacities, torch::kFloat32, "opacities", dev);
  Tensor bg = prep(bg_color, torch::kFloat32, "bg_color", dev);
  if (bg.numel() < 3) throw Exception("bg_color must have 3 entries");
  Tensor sh = shs.numel() > 0 ? prep(shs, torch::kFloat32, "shs", dev) : shs;
  const int M = (shs.numel() > 0 && shs.dim() >= 2) ? (int)shs.size(1) : 0;
  if (P > 0 && (sg.numel() < P || op.numel() < P)) throw Exception("sigma / opacities need number_of_points entries");
  if (P > 0 && shs.numel() > 0 && (shs.dim() != 3 || shs.size(0) != P || shs.size(2) != 3))
    throw Exception("shs must be [number_of_points, M, 3]");
  Tensor ex;
  int ex_len = 0;
  if (extra.defined() && extra.numel() > 0) {
    Tensor e = extra.is_cuda() ? extra : extra.to(torch::Device(torch::kCUDA, dev));
    ex = e.contiguous().view(torch::kUInt8).view(torch::kFloat32);
    ex_len = (int)ex.numel();
  }
  const int num_vert = (int)(verts.numel() / 3);

  void *stream = cur_stream(dev);
  // rebuild_gas: 0 build, 1 refit, 2 reuse (tracer.cpp:181-195)
  if (rebuild_gas_opt == 0 || rebuild_gas_opt == 1) {
    fwdState.d_vertices = verts.data_ptr<float>();
    fwdState.d_cumsum = cums.data_ptr<int>();
    if (rebuild_gas_opt == 0) fwdState.indexbuf_tensor = cums;
    build_tri_accel(fwdState, (uint32_t)num_vert, (uint32_t)P, rebuild_gas_opt == 0);
  } else if (rebuild_gas_opt == 3) {  // this build's addition: refit, rebuild when the SAH cost has degraded
    fwdState.d_vertices = verts.data_ptr<float>();
    fwdState.d_cumsum = cums.data_ptr<int>();
    check_tt(tt_build(tr_, fwdState.d_vertices, num_vert, fwdState.d_cumsum, P, TT_REFIT_AUTO, stream),
             "acceleration structure refit (automatic rebuild)");
  } else if (rebuild_gas_opt != 2) {
    throw Exception("rebuild_gas must be 0 (build), 1 (refit), 2 (reuse) or 3 (refit, rebuild when degraded)");
  } else if (tt_num_primitives(tr_) < 0) {
    throw Exception("rebuild_gas = 2 (reuse) but no acceleration structure has been built");
  }

  auto fopts = torch::TensorOptions().dtype(torch::kFloat32).device(torch::kCUDA, dev);
  Tensor out_color = torch::zeros({3, H, W}, fopts);
  Tensor out_others = torch::zeros({7, H, W}, fopts);
  Tensor max_blending = torch::zeros({P}, fopts);
  Tensor scal = prep(scaling, torch::kFloat32, "scaling", dev);
  Tensor dens = prep(density_factor, torch::kFloat32, "density_factor", dev);
  if (P > 0 && (scal.numel() < P || dens.numel() < P)) throw Exception("scaling / density_factor need P entries");

  tt_fwd_args a{};
  a.width = W;
  a.height = H;
  a.ray_o_dev = ro.data_ptr<float>();
  a.ray_d_dev = rd.data_ptr<float>();
  a.vertices_dev = verts.data_ptr<float>();
  a.num_vertices = num_vert;
  a.npts_dev = npts.data_ptr<int>();
  a.cumsum_dev = cums.data_ptr<int>();
  a.P = P;
  a.shs_dev = M > 0 ? sh.data_ptr<float>() : nullptr;
  a.M = M;
  a.sh_degree = sh_degree;
  a.opacities_dev = op.data_ptr<float>();
  a.sigma_dev = sg.data_ptr<float>();
  a.bg_dev = bg.data_ptr<float>();
  a.extra_dev = ex_len > 0 ? ex.data_ptr<float>() : nullptr;
  a.extra_len = ex_len;
  a.variant = variant_;
  a.k = k_;
  a.chunk_mode = chunk_mode_;
  a.cutoff = cutoff_;
  a.near_t = 0.2f;
  a.far_t = 1000.0f;
  a.color_dev = out_color.data_ptr<float>();
  a.others_dev = out_others.data_ptr<float>();
  a.max_blending_dev = max_blending.data_ptr<float>();
  a.density_dev = dens.data_ptr<float>();
  a.scaling_dev = scal.data_ptr<float>();
  Tensor hits, hit_counts, hit_colors;
  if (hit_cache_ > 0 && store_hits_ && variant_ == TT_VARIANT_FORWARD && (long long)H * W > 0) {
    auto iopts = torch::TensorOptions().dtype(torch::kInt32).device(torch::kCUDA, dev);
    hits = torch::empty({hit_cache_, (long long)H * W, 2}, iopts);
    hit_counts = torch::empty({(long long)H * W}, iopts);
    hit_colors = torch::empty({hit_cache_, (long long)H * W, 3}, fopts);
    a.hitlist_dev = reinterpret_cast<unsigned int *>(hits.data_ptr<int>());
    a.hitlist_count_dev = hit_counts.data_ptr<int>();
    a.hitlist_color_dev = hit_colors.data_ptr<float>();
    a.hitlist_cap = hit_cache_;
  }

  hipEvent_t start = nullptr, stop = nullptr;
  float elapsed_ms = 0.0f;
  if (debug) {
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess)
      throw Exception("hipEventCreate failed");
    (void)hipEventRecord(start, static_cast<hipStream_t>(stream));
  }
  check_tt(tt_trace_fwd(tr_, &a, stream), "trace_forward");
  if (debug) {
    (void)hipEventRecord(stop, static_cast<hipStream_t>(stream));
    (void)hipEventSynchronize(stop);
    (void)hipEventElapsedTime(&elapsed_ms, start, stop);
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
  }
  // in/out statistics: write back if the caller's tensors were not contiguous
  if (scal.data_ptr() != scaling.data_ptr()) scaling.copy_(scal.view(scaling.sizes()));
  if (dens.data_ptr() != density_factor.data_ptr()) density_factor.copy_(dens.view(density_factor.sizes()));

  Tensor gpu_time = torch::zeros({1}, fopts);
  if (debug) gpu_time.fill_(elapsed_ms);
  FwdTraceResult res;
  res.data = std::make_tuple(out_color, out_others, gpu_time, scaling, density_factor, max_blending);
  res.hits = hits;
  res.hit_counts = hit_counts;
  res.hit_colors = hit_colors;
  return res;
}

BwdTraceResult OptixTracer::trace_bwd(const Tensor &ray_o, const Tensor &ray_d, const Tensor &background,
                                      const Tensor &triangles_points, const Tensor &sigma,
                                      const torch::Tensor &num_points_per_triangle,
                                      const torch::Tensor &cumsum_of_points_per_triangle, const torch::Tensor &colors,
                                      const torch::Tensor &others, const int number_of_points, const Tensor &shs,
                                      const int sh_degree, const Tensor &opacities, const Tensor &dL_dpixels,
                                      const Tensor &dL_dothers, const Tensor &scaling, const bool debug,
                                      const c10::optional<Tensor> &hits, const c10::optional<Tensor> &hit_counts,
                                      const c10::optional<Tensor> &hit_colors) {
  (void)background;
  (void)scaling;
  (void)debug;
  const int dev = device_;
  c10::hip::HIPGuard guard(dev);
  if (ray_o.dim() != 3 || ray_o.size(2) != 3) throw Exception("ray_o must be [H, W, 3]");
  const int H = (int)ray_o.size(0), W = (int)ray_o.size(1);
  const int P = number_of_points;
  if (P < 0) throw Exception("number_of_points must be non-negative");
  Tensor ro = prep(ray_o, torch::kFloat32, "ray_o", dev);
  Tensor rd = prep(ray_d, torch::kFloat32, "ray_d", dev);
  Tensor verts = prep(triangles_points, torch::kFloat32, "triangles_points", dev);
  Tensor npts = prep(num_points_per_triangle, torch::kInt32, "num_points_per_triangle", dev);
  Tensor cums = prep(cumsum_of_points_per_triangle, torch::kInt32, "cumsum_of_points_per_triangle", dev);
  Tensor sg = prep(sigma, torch::kFloat32, "sigma", dev);
  Tensor op = prep(opacities, torch::kFloat32, "opacities", dev);
  Tensor col = prep(colors, torch::kFloat32, "colors", dev);
  Tensor oth = prep(others, torch::kFloat32, "others", dev);
  Tensor gp = prep(dL_dpixels, torch::kFloat32, "dL_dpixels", dev);
  Tensor go = prep(dL_dothers, torch::kFloat32, "dL_dothers", dev);
  if (col.numel() != 3ll * H * W || gp.numel() != 3ll * H * W) throw Exception("colors / dL_dpixels must be [3, H, W]");
  if (oth.numel() != 7ll * H * W || go.numel() != 7ll * H * W) throw Exception("others / dL_dothers must be [7, H, W]");
  Tensor sh = shs.numel() > 0 ? prep(shs, torch::kFloat32, "shs", dev) : shs;
  const int M = (shs.numel() > 0 && shs.dim() >= 2) ? (int)shs.size(1) : 0;

  // [npts[P-1] + cumsum[P-1], 3] (tracer.cpp:327-333); one device-to-host read
  int total_nb_points = 0;
  if (P > 0) total_nb_points = (npts[P - 1] + cums[P - 1]).item<int>();
  auto vopts = triangles_points.options();
  Tensor dL_dtriangle = torch::zeros({total_nb_points, 3}, vopts);
  Tensor dL_dsigma = torch::zeros_like(sigma);
  Tensor dL_dopacity = torch::zeros_like(opacities);
  Tensor dL_dsh = torch::zeros({P, M, 3}, vopts);
  if (P > 0 && (!dL_dsigma.is_contiguous() || !dL_dopacity.is_contiguous()))
    throw Exception("sigma / opacities must be contiguous");

  tt_bwd_args a{};
  a.width = W;
  a.height = H;
  a.ray_o_dev = ro.data_ptr<float>();
  a.ray_d_dev = rd.data_ptr<float>();
  a.vertices_dev = verts.data_ptr<float>();
  a.num_vertices = (int)(verts.numel() / 3);
  a.npts_dev = npts.data_ptr<int>();
  a.cumsum_dev = cums.data_ptr<int>();
  a.P = P;
  a.shs_dev = M > 0 ? sh.data_ptr<float>() : nullptr;
  a.M = M;
  a.sh_degree = sh_degree;
  a.opacities_dev = op.data_ptr<float>();
  a.sigma_dev = sg.data_ptr<float>();
  a.color_dev = col.data_ptr<float>();
  a.others_dev = oth.data_ptr<float>();
  a.dL_dpix_dev = gp.data_ptr<float>();
  a.dL_dothers_dev = go.data_ptr<float>();
  a.k = k_;
  a.chunk_mode = chunk_mode_;
  a.cutoff = cutoff_;
  a.near_t = 0.2f;
  a.far_t = 1000.0f;
  a.zero_grads = 0;
  a.dL_dvertices_rows = total_nb_points;
  a.dL_dvertices_dev = dL_dtriangle.data_ptr<float>();
  a.dL_dsigma_dev = dL_dsigma.data_ptr<float>();
  a.dL_dopacity_dev = dL_dopacity.data_ptr<float>();
  a.dL_dsh_dev = dL_dsh.data_ptr<float>();
  if (hits.has_value() && hits->defined() && hit_counts.has_value() && hit_counts->defined()) {
    const Tensor &hl = *hits, &hc = *hit_counts;
    if (hl.dim() != 3 || hl.size(1) != (long long)H * W || hl.size(2) != 2 || hl.scalar_type() != torch::kInt32 ||
        !hl.is_contiguous() || hc.numel() != (long long)H * W || hc.scalar_type() != torch::kInt32 ||
        !hc.is_contiguous() || hl.device() != ro.device() || hc.device() != ro.device())
      throw Exception("hits / hit_counts must be the forward's int32 [cap, H*W, 2] / [H*W] tensors");
    a.hitlist_dev = reinterpret_cast<const unsigned int *>(hl.data_ptr<int>());
    a.hitlist_count_dev = hc.data_ptr<int>();
    a.hitlist_cap = (int)hl.size(0);
    if (hit_colors.has_value() && hit_colors->defined()) {
      const Tensor &hk = *hit_colors;
      if (hk.dim() != 3 || hk.size(0) != hl.size(0) || hk.size(1) != hl.size(1) || hk.size(2) != 3 ||
          hk.scalar_type() != torch::kFloat32 || !hk.is_contiguous() || hk.device() != ro.device())
        throw Exception("hit_colors must be the forward's float32 [cap, H*W, 3] tensor");
      a.hitlist_color_dev = hk.data_ptr<float>();
    }
    // the forward's hit list is only valid for the scene that forward traced, so handing it
    // over already vouches for unchanged vertices and sigma: the backward may reuse that
    // forward's window records (the library still checks the pointers)
    a.same_scene_as_forward = 1;
  }
  check_tt(tt_trace_bwd(tr_, &a, cur_stream(dev)), "trace_backward");
  BwdTraceResult res;
  res.data = std::make_tuple(dL_dtriangle, dL_dsigma, dL_dopacity, dL_dsh);
  return res;
}
