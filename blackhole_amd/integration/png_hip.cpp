// png_hip.cpp - drop-in for the reference's PNG / APNG writer.
//
// Defines img::write_png and img::write_apng (src/image.h:45-50) over the
// GPU-assisted writer of libbhx.so (include/bhx_render.hpp): the file bytes
// are the reference's (src/image.cpp:193-778).  The reference's own
// definitions in src/image.cpp are set aside at build time by compiling that
// file with -Dwrite_png=write_png_cpu -Dwrite_apng=write_apng_cpu (see
// oracle/Makefile, target cli); every caller (src/main.cpp's stills and
// animations) is unchanged.
#include "image.h"

#include "bhx_render.hpp"

namespace img {

bool write_png(const std::string& path, const std::vector<uint8_t>& rgb, int w, int h) {
    return bhx::write_png(path, rgb, w, h);
}

bool write_apng(const std::string& path, const std::vector<std::vector<uint8_t>>& frames, int w, int h, int fps,
                int loops) {
    return bhx::write_apng(path, frames, w, h, fps, loops);
}

}  // namespace img
