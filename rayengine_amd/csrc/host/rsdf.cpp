// .rsdf header checks, shared by the loader and re_rsdf_info.
#include "rsdf.hpp"

#include <sys/types.h>

namespace re {

re_status rsdf_open(const char* path, FILE** out, int32_t dim[3], float spacing[3], const char** why) {
  *out = nullptr;
  *why = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    *why = "cannot open the file";
    return RE_ERR_IO;
  }
  const bool ok = std::fread(dim, sizeof(int32_t), 3, f) == 3 && std::fread(spacing, sizeof(float), 3, f) == 3;
  if (!ok) *why = "short header";
  else if (!valid_volume_dims(dim, spacing)) *why = "header dims outside [2,4096] or spacing not > 0";
  int64_t fsize = 0, want = 0;
  if (!*why) {
    want = 24 + 8 * (int64_t)dim[0] * dim[1] * dim[2];
    if (fseeko(f, 0, SEEK_END) != 0) *why = "cannot seek";
    else fsize = (int64_t)ftello(f);
  }
  if (!*why && fsize < want) *why = "file shorter than its header says";
  if (!*why && fsize > want) *why = "file longer than its header says";
  if (!*why && fseeko(f, 24, SEEK_SET) != 0) *why = "cannot seek";
  if (*why) {
    std::fclose(f);
    return RE_ERR_IO;
  }
  *out = f;
  return RE_OK;
}

}  // namespace re

extern "C" RE_API re_status re_rsdf_info(const char* path, int32_t dim[3], float spacing[3]) {
  if (!path || !dim || !spacing) return re::fail(RE_ERR_INVALID, "re_rsdf_info: null argument");
  FILE* f = nullptr;
  const char* why = nullptr;
  const re_status st = re::rsdf_open(path, &f, dim, spacing, &why);
  if (f) std::fclose(f);
  if (st) return re::fail(st, "re_rsdf_info: %s: %s", why, path);
  return RE_OK;
}
