// Pinhole camera rays on the GPU: the ray-generation step that feeds the tracer
// (the reference leaves it to the Python caller, diff_tritracer/__init__.py:54-139
// takes ray_o / ray_d [H, W, 3] as inputs). The camera basis comes from the caller;
// per pixel the direction is formed in double in a fixed order and rounded once
// to float, so the rays equal the host generator's bit for bit.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/tritracer_hip.h"
#include "tt_device.h"
#include "tt_internal.h"

namespace {

struct Cam {
  double c[tt::kCamDoubles];  // eye, forward, right, up, tan_x, tan_y
  int W, H;
};

__global__ void k_camera_rays(Cam c, float *ro, float *rd) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long long)c.W * c.H) return;
  float o[3], d[3];
  tt::pinhole_ray(c.c, c.W, c.H, (int)(p % c.W), (int)(p / c.W), o, d);
  for (int i = 0; i < 3; i++) {
    rd[3 * p + i] = d[i];
    ro[3 * p + i] = o[i];
  }
}

}  // namespace

extern "C" int tt_camera_rays(int width, int height, const double *eye, const double *forward, const double *right,
                              const double *up, double tan_x, double tan_y, float *ray_o_dev, float *ray_d_dev,
                              void *stream) {
  if (width < 0 || height < 0 || (long long)width * height >= (1ll << 31)) {
    tt::set_last_error("image size out of range");
    return TT_ERR_INVALID_ARG;
  }
  if ((long long)width * height == 0) return TT_OK;
  if (!eye || !forward || !right || !up || !ray_o_dev || !ray_d_dev) {
    tt::set_last_error("NULL argument");
    return TT_ERR_INVALID_ARG;
  }
  Cam c;
  for (int i = 0; i < 3; i++) c.c[i] = eye[i], c.c[3 + i] = forward[i], c.c[6 + i] = right[i], c.c[9 + i] = up[i];
  c.c[12] = tan_x;
  c.c[13] = tan_y;
  c.W = width;
  c.H = height;
  const long long n = (long long)width * height;
  hipLaunchKernelGGL(k_camera_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     c, ray_o_dev, ray_d_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    tt::set_last_error(std::string("k_camera_rays: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  return TT_OK;
}
