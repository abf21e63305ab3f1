// Driver for the CPU sanitizer run of the host C++ (TEST INFRASTRUCTURE): built
// with -fsanitize=address,undefined together with the scene library's sources
// and the oracle, it feeds them the inputs named on its command line; any
// out-of-bounds access, use after free, leak or undefined behaviour aborts it.
//
//   sanitize_driver write <config> <path>    write a workload as a .scene file
//   sanitize_driver scene <path>...          dfs_read_scene each file (refusal is fine)
//   sanitize_driver obj <path>...            load each OBJ into a builder and compile
//   sanitize_driver oracle <config> <w> <h>  render one pass of a workload crop with the oracle
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/dayflower_scene.h"

extern "C" {
// the oracle's frame and counters (oracle/df_oracle.cpp)
typedef struct oracle_frame {
  int32_t width, height;
  float* film;
  int32_t* film_samples;
  float* image_float;
  uint8_t* image_bgra;
  int32_t* primitive_offsets;
  int64_t* seeds;
} oracle_frame;
typedef struct oracle_counters {
  uint64_t rays, samples, box_tests, tri_tests, prim_tests, sphere_tests, terrain_steps, sky_evals;
} oracle_counters;
int oracle_render(const df_compiled_scene* scene, const df_sky* sky, const df_params* params, oracle_frame* frame,
                  int passes, int clear_first, int threads, int64_t pixel_begin, int64_t pixel_end,
                  oracle_counters* counters);
}

static int mode_write(int config, const char* path) {
  dfs_scene* s = nullptr;
  if (dfs_build_config(config, 0, &s) != DF_OK) return 2;
  const int rc = dfs_write_scene(s, path);
  dfs_free(s);
  return rc == DF_OK ? 0 : 3;
}

static int mode_scene(int n, char** paths) {
  for (int i = 0; i < n; ++i) {
    dfs_scene* s = nullptr;
    const int rc = dfs_read_scene(paths[i], &s);
    if (rc == DF_OK) {
      df_compiled_scene a;
      dfs_stats st;
      if (dfs_get_compiled(s, &a) != DF_OK || dfs_get_stats(s, &st) != DF_OK) return 4;
      dfs_free(s);
    }
    std::printf("scene %s %d\n", paths[i], rc);
  }
  return 0;
}

static int mode_obj(int n, char** paths) {
  for (int i = 0; i < n; ++i) {
    dfs_builder* b = nullptr;
    if (dfs_builder_create("sanitize", &b) != DF_OK) return 5;
    int32_t prims = 0;
    int rc = dfs_builder_load_obj(b, paths[i], nullptr, nullptr, -1, nullptr, &prims);
    if (rc == DF_OK) {
      const float plane_a[3] = {0, 0, 0}, plane_b[3] = {0, 0, 1}, plane_c[3] = {1, 0, 0};
      const float grey[3] = {0.5f, 0.5f, 0.5f}, black[3] = {0.0f, 0.0f, 0.0f};
      int32_t ta = 0, te = 0, surf = 0;
      rc = dfs_builder_constant_texture(b, grey, &ta);
      if (rc == DF_OK) rc = dfs_builder_constant_texture(b, black, &te);
      if (rc == DF_OK) rc = dfs_builder_surface(b, 3, ta, te, te, 0.0f, 0.0f, &surf);
      if (rc == DF_OK) rc = dfs_builder_plane(b, plane_a, plane_b, plane_c, surf, nullptr);
      dfs_scene* s = nullptr;
      if (rc == DF_OK && dfs_builder_compile(b, &s) == DF_OK) dfs_free(s);
    } else {
      char msg[256];
      dfs_builder_last_error(b, msg, sizeof msg);
    }
    dfs_builder_free(b);
    std::printf("obj %s %d %d\n", paths[i], rc, prims);
  }
  return 0;
}

static int mode_oracle(int config, int w, int h) {
  dfs_scene* s = nullptr;
  if (dfs_build_config(config, 0, &s) != DF_OK) return 6;
  df_compiled_scene a;
  df_sky sky;
  df_params p;
  int32_t sw, sh, spp;
  if (dfs_get_compiled(s, &a) != DF_OK || dfs_get_sky(s, &sky) != DF_OK ||
      dfs_get_render_config(s, &sw, &sh, &spp, &p) != DF_OK)
    return 7;
  std::vector<float> cam(a.camera, a.camera + 19);
  cam[16] = static_cast<float>(w);
  cam[17] = static_cast<float>(h);
  a.camera = cam.data();
  const size_t np = static_cast<size_t>(w) * h;
  std::vector<float> film(np * 3), imgf(np * 3);
  std::vector<int32_t> samples(np), prims(np, -1);
  std::vector<uint8_t> bgra(np * 4);
  std::vector<int64_t> seeds(np);
  for (size_t i = 0; i < np; ++i) seeds[i] = static_cast<int64_t>((i * 0x9E3779B97F4A7C15ull) & ((1ull << 48) - 1));
  oracle_frame f{w, h, film.data(), samples.data(), imgf.data(), bgra.data(), prims.data(), seeds.data()};
  oracle_counters c{};
  for (int renderer = 1; renderer <= 6; ++renderer) {
    p.renderer_type = renderer;
    if (oracle_render(&a, &sky, &p, &f, 1, 1, 2, 0, static_cast<int64_t>(np), &c) != 0) return 8;
  }
  dfs_free(s);
  std::printf("oracle %d %dx%d rays %llu\n", config, w, h, static_cast<unsigned long long>(c.rays));
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) return 1;
  if (!std::strcmp(argv[1], "write") && argc == 4) return mode_write(std::atoi(argv[2]), argv[3]);
  if (!std::strcmp(argv[1], "scene")) return mode_scene(argc - 2, argv + 2);
  if (!std::strcmp(argv[1], "obj")) return mode_obj(argc - 2, argv + 2);
  if (!std::strcmp(argv[1], "oracle") && argc == 5)
    return mode_oracle(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
  return 1;
}
