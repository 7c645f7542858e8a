// .rsdf volume files (Grid's on-disk form, RayEngine/src/World/Grid.cpp:23-64
// and save_map / load_map, RayEngine/src/cuda-src/renderer.cu:838-947):
// a 24-byte header of int32 dims and float spacing, then dim.x * dim.y *
// dim.z (distance, material) float pairs, x fastest.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>

#include "errors.hpp"

namespace re {

// Volume dims in [2, 4096] and finite spacing > 0 on every axis.
inline bool valid_volume_dims(const int32_t* dim, const float* spacing) {
  for (int a = 0; a < 3; ++a) {
    if (dim[a] < 2 || dim[a] > 4096) return false;
    if (!(spacing[a] > 0.0f) || !std::isfinite(spacing[a])) return false;
  }
  return true;
}

// Opens an .rsdf file and checks it before anything of the size its header
// claims is allocated: the header's dims and spacing valid and the file
// exactly 24 + 8 * dim.x * dim.y * dim.z bytes long.  RE_OK with the file
// positioned at the first pair, or RE_ERR_IO with *why set and no file open.
re_status rsdf_open(const char* path, FILE** out, int32_t dim[3], float spacing[3], const char** why);

}  // namespace re
