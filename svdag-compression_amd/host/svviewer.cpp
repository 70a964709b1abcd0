// svviewer -- headless form of the reference viewer (src/svviewer/main.cpp):
//   svviewer file.{svdag,ussvdag,ssvdag,esvdag} [--width W] [--height H]
//            [--frames N] [--seed S] [--no-shadows] [--no-beam] [--image out.ppm]
//            [--mode M] [--norm-samples K] [--select NODE] [--no-random-colors]
//            [--no-freq-colors] [--ao RAYS]
// Loads the file (EncodedOctree::load), traces frames with the octree-DDA
// kernel from a seeded orbit camera (renderer defaults: FOV 60, maxIters 300,
// pixel tolerance 1, drawLevel = levels, light at the box max) and prints
// Rays/s like the viewer's window title (svviewer/main.cpp:166-183).
// --image writes the last frame as the viewer draws it (svd_render: render
// mode M = viewerRenderMode 0 iterations / 1 depth / 2 levels / 3 pretty, K
// normal samples, the selected node, random and frequency colours; the
// framebuffer's clamp to [0, 1]); with --ao it writes the AO pass's visibility
// of RAYS hemisphere rays instead (svd_trace_ao).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "symvox_amd.hpp"

namespace {

uint64_t splitmix(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void invert4(const double m[16], double inv[16]) {   // row-major general inverse (Gauss-Jordan)
  double a[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 8; ++c) a[r][c] = c < 4 ? m[r * 4 + c] : (c - 4 == r ? 1.0 : 0.0);
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r) if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[p][k]);
    const double d = a[c][c];
    for (int k = 0; k < 8; ++k) a[c][k] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) inv[r * 4 + c] = a[r][c + 4];
}

// Same camera as the Python workloads.orbit_camera (seeded yaw, pitch 20-50 deg,
// radius 1.5 x half diagonal of the box, looking at the box centre).
void orbit_camera(uint64_t seed, const svd::aabox3f& box, uint32_t w, uint32_t h, float viewInv[16], float projInv[16]) {
  uint64_t s = seed;
  const double u0 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
  const double u1 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
  const double yaw = 2 * M_PI * u0, pitch = (20.0 + 30.0 * u1) * M_PI / 180.0;
  double c[3], diag = 0;
  for (int i = 0; i < 3; ++i) { c[i] = 0.5 * ((double)box[0][i] + box[1][i]); diag += std::pow((double)box[1][i] - box[0][i], 2); }
  const double radius = 1.5 * std::sqrt(diag) / 2.0;
  const double eye[3] = {c[0] + radius * std::cos(pitch) * std::cos(yaw), c[1] + radius * std::sin(pitch),
                         c[2] + radius * std::cos(pitch) * std::sin(yaw)};
  double f[3] = {c[0] - eye[0], c[1] - eye[1], c[2] - eye[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double& x : f) x /= fl;
  const double up[3] = {0, 1, 0};
  double sv[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double sl = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
  for (double& x : sv) x /= sl;
  const double uv[3] = {sv[1] * f[2] - sv[2] * f[1], sv[2] * f[0] - sv[0] * f[2], sv[0] * f[1] - sv[1] * f[0]};
  double V[16] = {sv[0], sv[1], sv[2], -(sv[0] * eye[0] + sv[1] * eye[1] + sv[2] * eye[2]),
                  uv[0], uv[1], uv[2], -(uv[0] * eye[0] + uv[1] * eye[1] + uv[2] * eye[2]),
                  -f[0], -f[1], -f[2], (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]),
                  0, 0, 0, 1};
  const double t = 1.0 / std::tan(M_PI / 6.0), near = 0.01, far = 100.0, aspect = double(w) / h;
  double P[16] = {t / aspect, 0, 0, 0, 0, t, 0, 0, 0, 0, (far + near) / (near - far), 2 * far * near / (near - far),
                  0, 0, -1, 0};
  double Vi[16], Pi[16];
  invert4(V, Vi);
  invert4(P, Pi);
  for (int r = 0; r < 4; ++r)
    for (int cc = 0; cc < 4; ++cc) { viewInv[cc * 4 + r] = (float)Vi[r * 4 + cc]; projInv[cc * 4 + r] = (float)Pi[r * 4 + cc]; }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    printf("Usage: svviewer file.{svdag,ussvdag,ssvdag,esvdag} [--width W] [--height H] [--frames N] [--seed S] "
           "[--no-shadows] [--no-beam] [--image out.ppm] [--mode M] [--norm-samples K] [--select NODE] "
           "[--no-random-colors] [--no-freq-colors] [--ao RAYS]\n");
    return 1;
  }
  std::string file = argv[1];
  uint32_t W = 1280, H = 720, frames = 10;
  uint64_t seed = 1;
  bool shadows = true, beam = true;   // the viewer's minimum-depth optimisation is on by default
  std::string image;
  // OctreeDDARenderer defaults (octree_dda_renderer.cpp:41-45, 73)
  svd_view_params view{0, 0, 0xFFFFFFFFu, 1, 1, 0};
  int aoRays = -1;
  for (int i = 2; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--width" && i + 1 < argc) W = atoi(argv[++i]);
    else if (a == "--height" && i + 1 < argc) H = atoi(argv[++i]);
    else if (a == "--frames" && i + 1 < argc) frames = atoi(argv[++i]);
    else if (a == "--seed" && i + 1 < argc) seed = strtoull(argv[++i], nullptr, 10);
    else if (a == "--no-shadows") shadows = false;
    else if (a == "--no-beam") beam = false;
    else if (a == "--image" && i + 1 < argc) image = argv[++i];
    else if (a == "--mode" && i + 1 < argc) view.render_mode = atoi(argv[++i]);
    else if (a == "--norm-samples" && i + 1 < argc) view.norm_samples = atoi(argv[++i]);
    else if (a == "--select" && i + 1 < argc) view.selected_voxel = (uint32_t)strtoul(argv[++i], nullptr, 10);
    else if (a == "--no-random-colors") view.random_colors = 0;
    else if (a == "--no-freq-colors") view.freq_colors = 0;
    else if (a == "--ao" && i + 1 < argc) aoRays = atoi(argv[++i]);
  }
  EncodedOctree* eo = nullptr;
  EncodedSVDAG svdag;
  EncodedUSSVDAG ussvdag;
  EncodedSSVDAG ssvdag;
  // svviewer/main.cpp:193-207: the extension selects the format
  if (file.size() > 8 && file.compare(file.size() - 8, 8, ".ussvdag") == 0) eo = &ussvdag;
  else if (file.size() > 7 && (file.compare(file.size() - 7, 7, ".ssvdag") == 0 || file.compare(file.size() - 7, 7, ".esvdag") == 0)) eo = &ssvdag;
  else if (file.size() > 6 && file.compare(file.size() - 6, 6, ".svdag") == 0) eo = &svdag;
  else { printf("Unsupported file type '%s'\n", file.c_str()); return 1; }
  if (!eo->load(file)) return 1;
  printf("* %s: %u levels, %zu nodes, %zu bytes of data\n", eo->getDescription().c_str(), eo->getNLevels(),
         eo->getNNodes(), eo->getDataSize());
  svd_trace_params p{};
  orbit_camera(seed, eo->getSceneBBox(), W, H, p.view_inv, p.proj_inv);
  p.width = W;
  p.height = H;
  p.max_iters = 300;
  p.draw_level = eo->getNLevels();
  // getProjectionFactor (octree_dda_renderer.cpp:503-507), pixel tolerance 1
  const float fovy = 60.0f * 3.14159265359f / 180.0f;
  p.projection_factor = (1.0f / (2.0f * std::tan(0.5f * fovy))) / (1.0f / (float)H);
  // getProjectionFactor(1, 8) of the depth pre-pass
  const float beamProjection = beam ? (1.0f / (2.0f * std::tan(0.5f * fovy))) / (1.0f / ((float)H / 8.0f)) : 0.f;
  const svd::aabox3f bb = eo->getSceneBBox();
  p.light_pos[0] = bb[1][0]; p.light_pos[1] = bb[1][1]; p.light_pos[2] = bb[1][2];
  p.shadows = shadows ? 1 : 0;
  double best = 1e30;
  for (uint32_t f = 0; f < frames; ++f) {
    HeadlessTraceResult r;
    if (!traceHeadless(*eo, p, r, nullptr, beamProjection)) return 1;
    best = std::min(best, r.kernelMs);
    printf("frame %u: %.3f ms, %lu primary + %lu shadow rays, %.1f Mrays/s (hits %lu)\n", f, r.kernelMs,
           (unsigned long)r.primaryRays, (unsigned long)r.shadowRays, r.mraysPerSec, (unsigned long)r.hits);
  }
  printf("Rays/s = %.1f M (best frame %.3f ms)\n", (double)W * H / (best * 1e-3) / 1e6, best);
  if (!image.empty()) {
    // the framebuffer stores clamp(c, 0, 1) in 8 bits; rows bottom-up in GL
    std::vector<float> px;
    double ms = 0;
    int ch = 3;
    if (aoRays >= 0) {
      if (!aoHeadless(*eo, p, (uint32_t)aoRays, px, &ms)) return 1;
      ch = 1;
      printf("AO pass (%d rays per pixel): %.3f ms\n", aoRays, ms);
    } else {
      if (!renderHeadless(*eo, p, view, beamProjection, px, &ms)) return 1;
      printf("viewer frame (mode %d, %d normal samples): %.3f ms\n", view.render_mode, view.norm_samples, ms);
    }
    FILE* fp = fopen(image.c_str(), "wb");
    if (!fp) { printf("ERROR: cannot write %s\n", image.c_str()); return 1; }
    fprintf(fp, "P6\n%u %u\n255\n", W, H);
    for (int y = (int)H - 1; y >= 0; --y)
      for (uint32_t x = 0; x < W; ++x) {
        uint8_t out[3];
        for (int k = 0; k < 3; ++k) {
          float v = px[((size_t)y * W + x) * ch + (ch == 3 ? k : 0)];
          v = v == v ? std::min(1.0f, std::max(0.0f, v)) : 0.0f;
          out[k] = (uint8_t)std::lrint(v * 255.0f);
        }
        fwrite(out, 1, 3, fp);
      }
    fclose(fp);
    printf("* wrote %s\n", image.c_str());
  }
  return 0;
}
